/*
 * plato_mi355x.h -- C ABI of the MI355X (gfx950) Tet4 SIMP elastostatics library.
 *
 * One handle holds one mesh, one isotropic material, one MSIMP penalty, its essential
 * and natural boundary conditions and the device data derived from them (block graph,
 * slot codes, LDS chunks).  All compute entry points take DEVICE pointers (HBM) and run
 * on the handle's HIP stream; set-up entry points take HOST pointers.  Every entry
 * point returns PA_OK (0) or a non-zero status; pa_last_error() gives the message of
 * the calling thread's last failure.  A handle is not thread-safe: one thread (and one
 * stream) per handle, as the reference managers are single-threaded on the host.
 *
 * Layouts follow the reference (src/PlatoTypes.hpp:15-24, LayoutRight):
 *   coords  f64[3*num_nodes]  x,y,z per node
 *   conn    i32[4*num_elems]  Tet4 vertices per element
 *   u, R    f64[3*num_nodes]  dof 3*n+i
 *   z       f64[num_nodes]    nodal density (control)
 *   Jacobian values f64[9*nnz_b], block (slot s, i, j) at 9*s+3*i+j
 *   gradient_z values f64[3*nnz_b], transposed 1x3 blocks, entry 3*s+i
 */
#ifndef PLATO_MI355X_H
#define PLATO_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum pa_status {
  PA_OK = 0,
  PA_ERR_ARG = 1,          /* bad argument (null pointer, size, index) */
  PA_ERR_HIP = 2,          /* HIP runtime error (allocation, launch)    */
  PA_ERR_MESH = 3,         /* mesh/graph inconsistent                   */
  PA_ERR_STATE = 4,        /* call order (e.g. compute before finalize) */
  PA_ERR_NOT_CONVERGED = 5,/* linear solver hit its iteration cap       */
  PA_ERR_UNSUPPORTED = 6   /* mesh outside the kernels' limits          */
};

typedef struct pa_problem* pa_handle;

/* Message of the calling thread's last failed call ("" if none). */
const char* pa_last_error(void);

/* ---------------------------------------------------------------- mesh utilities (host)
 * Box [0,lx]x[0,ly]x[0,lz] with nx*ny*nz hexes split into 6 Tet4 each, numbered like
 * Omega_h::build_box (used by the reference tests: unit_tests/PlatoTestHelpers.hpp:206-250).
 * Call with coords==NULL/conn==NULL to get the sizes, then again with buffers of
 * 3*num_nodes doubles and 4*num_elems ints. */
int pa_build_box(int nx, int ny, int nz, double lx, double ly, double lz, int* num_nodes,
                 int* num_elems, double* coords, int32_t* conn);

/* Node set of a coordinate plane (mesh "node set", src/EssentialBCs.hpp:185-218 consumes
 * it).  out==NULL returns the count only. */
int pa_nodes_on_plane(int num_nodes, const double* coords, int axis, double value, double tol,
                      int32_t* out, int* count);

/* Side set of a coordinate plane as face vertex triples plus the adjacent element
 * (what src/SurfaceLoadIntegral.hpp:100-157 walks).  NULL outputs return the count. */
int pa_faces_on_plane(int num_elems, const int32_t* conn, const double* coords, int axis,
                      double value, double tol, int32_t* face_verts, int32_t* face_elems,
                      int* count);

/* Block graph of a mesh on the host: node star plus diagonal, ascending columns
 * (CreateBlockMatrix, src/ImplicitFunctors.hpp:735-808).  colidx==NULL returns
 * rowmap (num_nodes+1 entries) and *nnz_blocks only. */
int pa_node_graph(int num_nodes, int num_elems, const int32_t* conn, int32_t* rowmap, int32_t* colidx,
                  int* nnz_blocks);

/* ---------------------------------------------------------------- problem set-up (host)
 * Mesh arrays are copied; the caller keeps ownership. */
int pa_create(pa_handle* h, int num_nodes, int num_elems, const double* coords,
              const int32_t* conn);
int pa_destroy(pa_handle h);

/* "Isotropic Linear Elastic" material: Youngs Modulus, Poissons Ratio
 * (src/IsotropicLinearElasticMaterial.cpp:87-127). */
int pa_set_material(pa_handle h, double youngs_modulus, double poissons_ratio);

/* MSIMP penalty: "Exponent" p, "Minimum Value" rho_min (src/Simp.hpp:44-50,130-134):
 * w = rho_min + (1-rho_min)*pow(mean(z_e), p).  Defaults 3 and 1e-9 as the reference. */
int pa_set_penalty(pa_handle h, double exponent, double minimum_value);
/* Penalty model of the stiffness weight w(rho) applied to the stress (ApplyWeighting,
 * src/ApplyWeighting.hpp:49-65); "Penalty Function" Type of the reference's input:
 *   0 SIMP       w = rmin + (1 - rmin) rho^p                      (src/Simp.hpp:130-134)
 *   1 RAMP       w = rmin + (1 - rmin) rho / (1 + p (1 - rho))    (src/Ramp.hpp)
 *   2 Heaviside  w = rmin + (1 - rmin) (0.5 (1 + sin(pi rho / (2 L))))^p on |rho| < L,
 *                rmin below -L, 1 above L                         (src/Heaviside.hpp)
 *   3 none       w = 1                                            (src/NoPenalty.hpp)
 * exponent = "Exponent", minimum_value = "Minimum Value", length = "Regularization
 * Length" (Heaviside only).  pa_set_penalty(h, p, rmin) is model 0. */
int pa_set_penalty_model(pa_handle h, int model, double exponent, double minimum_value, double length);

/* Essential BC on one dof of a node set (src/EssentialBCs.hpp:185-218), value scaled like
 * the reference's step scale in applyBlockConstraints. */
int pa_add_essential_bc(pa_handle h, int num_nodes, const int32_t* nodes, int dof, double value);

/* Uniform traction on a side set (src/NaturalBC.hpp:186-228, SurfaceLoadIntegral):
 * each face vertex receives area/3 * traction in f, and R = K u - f. */
int pa_add_traction(pa_handle h, int num_faces, const int32_t* face_verts,
                    const int32_t* face_elems, const double traction[3]);

/* Optional caller block graph (rowmap i32[num_nodes+1], colidx i32[nnz_b]); without it
 * the node star plus diagonal is built with ascending columns
 * (CreateBlockMatrix, src/ImplicitFunctors.hpp:735-808).  Assembly writes values in the
 * order of whichever graph the handle holds, so a reference graph is reproduced exactly. */
int pa_set_graph(pa_handle h, const int32_t* rowmap, const int32_t* colidx);

/* Builds the device data (graph, slot codes, chunks, loads, BC lists).  Must be called
 * once after the set-up calls and before any compute call. */
int pa_finalize(pa_handle h);

/* HIP stream for all later launches (hipStream_t passed as void*; NULL = default). */
int pa_set_stream(pa_handle h, void* hip_stream);

int pa_graph_size(pa_handle h, int* num_rows, int* nnz_blocks);
/* Copies the graph out; on_device selects device (1) or host (0) destination buffers. */
int pa_copy_graph(pa_handle h, int32_t* rowmap, int32_t* colidx, int on_device);

/* ------------------------------------------------ Plato::Elliptic::VectorFunction (device)
 * value: R(u,z) = K(z) u - f          (src/elliptic/VectorFunction.hpp:231-313)
 * gradient_u: K(z) into values, optionally R in the same pass (R may be NULL)
 *                                      (src/elliptic/VectorFunction.hpp:505-594; the
 *                                      gradient_u_T of :412 is the same matrix here)
 * gradient_z: dR/dz as transposed 1x3 blocks on the same graph
 *                                      (src/elliptic/VectorFunction.hpp:598-689) */
int pa_vf_value(pa_handle h, const double* u, const double* z, double* R);
int pa_vf_gradient_u(pa_handle h, const double* u, const double* z, double* values, double* R);
int pa_vf_gradient_z(pa_handle h, const double* u, const double* z, double* values_t);
/* gradient_x: dR/dX as transposed 3x3 blocks on the same graph (9*nnz_b): block (config
 * node b, residual node a), entry (k, i) = dR_{a,i}/dX_{b,k}, including the traction
 * load's dependence on its faces (src/elliptic/VectorFunction.hpp:317-408, assembled
 * by assembleTransposeJacobian, src/Assembly.hpp:1422-1485) */
int pa_vf_gradient_x(pa_handle h, const double* u, const double* z, double* values_t);

/* ---------------------------------- Plato::Elliptic::PhysicsScalarFunction, criterion
 * "Internal Elastic Energy" (src/elliptic/InternalElasticEnergy.hpp:100-158):
 * value C = sum_e 0.5*vol*sigma_w:eps   (PhysicsScalarFunction.hpp:225-283), value to host
 * gradient_u = K(z) u                   (PhysicsScalarFunction.hpp:365-…)
 * gradient_z = partial dC/dz (nodal)    (PhysicsScalarFunction.hpp:432-495) */
int pa_sf_value(pa_handle h, const double* u, const double* z, double* value);
int pa_sf_gradient_u(pa_handle h, const double* u, const double* z, double* g);
int pa_sf_gradient_z(pa_handle h, const double* u, const double* z, double* g, double* value);
/* Geometric "Volume" criterion (src/geometric/Volume.hpp:49-96) with its own penalty
 * function (model as pa_set_penalty_model): V = sum_e vol_e w(mean z_e); dV/dz (N) and
 * dV/dX (3N) as GeometryScalarFunction gradient_z / gradient_x
 * (src/geometric/GeometryScalarFunction.cpp).  value (host), dvdz, dvdx may be NULL. */
int pa_volume(pa_handle h, const double* z, int model, double exponent, double minimum_value, double length,
              double* value, double* dvdz, double* dvdx);
/* gradient_x = partial dC/dX at fixed u and z (3N)
 *                                       (PhysicsScalarFunction.hpp:293-355) */
int pa_sf_gradient_x(pa_handle h, const double* u, const double* z, double* g);

/* ------------------------------------------------------------ Problem / solver (device) */
typedef struct {
  double tolerance;   /* relative residual ||b-Ax||/||b|| to stop at                 */
  int max_iters;      /* iteration cap (PA_ERR_NOT_CONVERGED when reached)           */
  int check_every;    /* host convergence check period (iterations; multigrid: <= 2) */
  int precision;      /* 0: FP64 block-CRS PCG; 1: FP64 PCG with an FP32 matrix-free operator */
  int use_x0;         /* 1: start from the contents of x (warm start), 0: from zero  */
  double inner_tolerance; /* precision 1: drop of the iterated residual between true FP64
                             residual replacements (clamped to [1e-6, 0.5]; 0 -> 0.1)     */
  int preconditioner;     /* 0: 3x3 block Jacobi; 1: geometric multigrid V-cycle (needs
                             pa_set_box; pa_solution and pa_stress_pnorm_gradient only)  */
} pa_solver_opts;

typedef struct {
  int iterations;       /* total inner iterations                                    */
  int outer_iterations; /* FP64 residual replacements (precision 1)                  */
  int converged;
  double rel_residual;  /* final true FP64 ||b-Ax||/||b||                            */
} pa_solver_stats;

/* applyBlockConstraints<3> (src/ApplyConstraints.hpp:11-79): BC rows/cols of `values`
 * zeroed with unit diagonal, rhs lifted and set to scale*value. */
int pa_apply_constraints(pa_handle h, double* values, double* rhs, double scale);

/* AbstractSolver::solve(A, x, b) (src/alg/PlatoAbstractSolver.hpp:19-27) for a
 * constrained block-CRS matrix on the handle's graph: Jacobi(3x3 block)-PCG in FP64. */
int pa_solve(pa_handle h, const double* values, double* x, const double* b,
             const pa_solver_opts* opts, pa_solver_stats* stats);

/* No counterpart in the reference.  pa_solve for callers that cannot allocate device
 * memory during the call (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same block-Jacobi PCG, same arguments and results, with its
 * workspace (four node vectors, the block-diagonal inverse, reduction partials and
 * scalars) in a fixed device array of the library instead of handle memory allocated on
 * first use.  Meshes of at most PA_NOALLOC_MAX_NODES nodes (more: PA_ERR_UNSUPPORTED);
 * the array (21 * PA_NOALLOC_MAX_NODES doubles plus 6160, about 22 MB) is held on every
 * GPU the library runs on until the process ends.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (other streams, other threads, other
 * handles) race and must be serialised by the caller. */
#define PA_NOALLOC_MAX_NODES 131072
int pa_solve_noalloc(pa_handle h, const double* values, double* x, const double* b,
                     const pa_solver_opts* opts, pa_solver_stats* stats);

/* Problem::solution (src/elliptic/Problem.hpp:259-311): u = 0, rhs = -R(0,z) = f,
 * K = gradient_u, constraints, solve, u += du.  `values` is caller scratch of 9*nnz_b
 * doubles (may be NULL for precision 1, which never forms K).
 * precision 1: Jacobi-PCG with FP64 vectors and reductions whose operator K p is applied
 * matrix-free in FP32; each time the iterated residual has dropped by inner_tolerance it
 * is replaced by the true FP64 residual f - K u of the FP64 element kernel (the search
 * direction is kept); convergence is judged on that true residual.  use_x0 = 1 (either
 * precision, either preconditioner) starts from the caller's u (warm start; its
 * constrained dofs are reset to their prescribed values). */
int pa_solution(pa_handle h, const double* z, double* u, double* values,
                const pa_solver_opts* opts, pa_solver_stats* stats);

/* No counterpart in the reference (its solves start from zero, src/elliptic/Problem.hpp:263).
 * Problem::solution in FP64 (assembled K, 3x3 block-Jacobi PCG) deflated by the k <= 16
 * device vectors W (column j at W + j*3N; e.g. the learned warm-start network's
 * unit-traction responses, or earlier solutions of an optimisation run): deflated PCG
 * (Saad, Yeung, Erhel, Guyomarc'h, SIAM J. Sci. Comput. 21(5), 2000) starts from the
 * Galerkin correction in span(W) (added to the caller's u when use_x0 = 1) and keeps
 * every search direction A-orthogonal to span(W).  Constrained dofs of W are ignored;
 * A-dependent vectors are dropped.  precision 0 and preconditioner 0 only; k = 0 is
 * pa_solution's FP64 block-Jacobi solve. */
int pa_solution_deflated(pa_handle h, const double* z, double* u, double* values, const double* W, int k,
                         const pa_solver_opts* opts, pa_solver_stats* stats);

/* Declares the mesh a structured box of nx*ny*nz hexes (6 Tet4 each) whose nodes lie on
 * the (nx+1)(ny+1)(nz+1) lattice spanning their bounding box, in any order and
 * numbering.  Enables preconditioner 1: a geometric multigrid hierarchy of boxes with
 * half the hexes per axis while the counts stay even, rediscretised with densities
 * homogenised through the SIMP stiffness weight, Chebyshev block-Jacobi smoothing and a
 * dense solve on the last level.  No counterpart in the reference, whose solver
 * interface defers to external algebraic multigrid (src/alg/AmgXLinearSolver.hpp,
 * src/alg/EpetraLinearSolver.cpp); PA_ERR_MESH when the nodes are not such a lattice. */
int pa_set_box(pa_handle h, int nx, int ny, int nz);

/* No counterpart in the reference (an introspection call for callers and tests): *active
 * = 1 when the handle assembles its Jacobian with the lattice kernel, which applies after
 * pa_set_box when every hex of the box carries the same six lattice tets, the nodes sit
 * on the uniform lattice (within 1e-13 spacings) and every graph row has at most 16
 * blocks; the values equal the general kernel's to rounding.  PA_BOX_ASSEMBLY=0 in the
 * environment at pa_set_box / pa_finalize keeps the general kernel. */
int pa_box_assembly(pa_handle h, int* active);

/* Damped block-Jacobi smoothing of the homogeneous problem, u <- u - omega D^-1 K(z) u
 * at the free dofs, `steps` times (D: 3x3 diagonal blocks of K).  A low-pass filter
 * adapted to K: modes with a large Rayleigh quotient are damped, smooth ones kept.  No
 * counterpart in the reference; used to denoise a learned initial guess before its
 * Galerkin scaling. */
int pa_smooth(pa_handle h, const double* z, double* u, int steps, double omega);

/* Problem::criterionGradient for the self-adjoint compliance
 * (src/elliptic/Problem.hpp:413-468, :437-441): dC/dz = -partial dC/dz at the solved u.
 * `value` (host, may be NULL) receives C. */
int pa_compliance_gradient(pa_handle h, const double* u, const double* z, double* dcdz,
                           double* value);

/* ------------------------------------------------- stress p-norm criterion and adjoint
 * F = ( sum_e vol_e * m_e^p )^(1/p) with
 *   measure 0: m_e = Voigt norm of the SIMP-weighted stress -- the reference
 *              "Stress P-Norm" (src/elliptic/StressPNorm.hpp:88-142, TensorPNorm
 *              src/TensorPNorm.hpp:28-45,286-318; p is truncated to an integer as there);
 *   measure 1: m_e = rho_e^q * vonMises(sigma0_e), the relaxed von Mises stress of the
 *              unpenalized stress (vonMises as src/VonMisesYieldFunction.hpp:53-74).
 * pa_stress_pnorm: value (host) and the partial derivatives dF/du (3N) and dF/dz (N),
 * either may be NULL (PhysicsScalarFunction::value / gradient_u / gradient_z). */
int pa_stress_pnorm(pa_handle h, const double* u, const double* z, int measure, double p, double q,
                    double* value, double* dfdu, double* dfdz);

/* PhysicsScalarFunction::gradient_x of the stress p-norm (src/elliptic/PhysicsScalarFunction.hpp:293-355
 * with StressPNorm, src/elliptic/StressPNorm.hpp:88-142): partial dF/dX at fixed u and z
 * (3N), including the cell volumes' dependence on the coordinates; value (host) may be NULL. */
int pa_stress_pnorm_gradient_x(pa_handle h, const double* u, const double* z, int measure, double p, double q,
                               double* dfdx, double* value);

/* (dR/dz)^T lambda per control node (VectorFunction::gradient_z applied transposed,
 * the MatrixTimesVectorPlusVector term of src/elliptic/Problem.hpp:437-466) without
 * forming dR/dz. */
int pa_dRdz_T_apply(pa_handle h, const double* u, const double* lambda, const double* z, double* out);

/* Plato::MatrixTimesVectorPlusVector (src/PlatoMathHelpers.hpp:58-138) on the handle's
 * block graph: y += alpha * A x.  block_rows 3: 3x3 blocks, 9 values per entry (the
 * gradient_u and gradient_x layouts, y has 3N entries); block_rows 1: 1x3 blocks, 3
 * values per entry (the transposed gradient_z layout, y has N entries).  x has 3N. */
int pa_matvec_plus_vector(pa_handle h, int block_rows, double alpha, const double* values, const double* x,
                          double* y);

/* Problem::criterionGradient, non-self-adjoint branch (src/elliptic/Problem.hpp:413-468):
 * K lambda = dF/du (homogeneous constraints; FP64 CRS PCG or mixed precision per opts),
 * dF/dz = dF/dz|_u - (dR/dz)^T lambda.  `lambda` (3N) receives the adjoint state. */
int pa_stress_pnorm_gradient(pa_handle h, const double* u, const double* z, int measure, double p, double q,
                             double* dfdz, double* lambda, double* value, const pa_solver_opts* opts,
                             pa_solver_stats* stats);

/* ------------------------------------------------------------ learned warm start
 * No counterpart in the reference (whose Problem::solution always starts from u = 0):
 * a 4-level 3-D convolutional U-Net on the node grid of a box mesh, trained by the
 * library's own BF16-MFMA forward/backward kernels, whose output warm-starts the solve
 * (pa_solution with opts.use_x0 = 1).  Tensors are device FP32, channels last:
 * input [batch][nx][ny][nz][8] — density, load (3), coordinates (3), SIMP weight;
 * output [batch][nx][ny][nz][16] — u (3), or u for three unit load cases (9), then
 * padding channels (trained towards 0).  Node grid
 * dimensions must be 1 mod 4.  Parameters: FP32 vector of pa_unet_num_params values
 * (per layer: W[27][cin][cout] then bias[cout]). */
typedef struct pa_unet_s* pa_unet;
const char* pa_unet_last_error(void);
int pa_unet_create(pa_unet* out, int nx, int ny, int nz, int base_channels, unsigned long long seed);
int pa_unet_destroy(pa_unet h);
int pa_unet_set_stream(pa_unet h, void* hip_stream);
int pa_unet_num_params(pa_unet h, size_t* n);
/* copy parameters out to dst and/or in from src (device pointers, either may be NULL) */
int pa_unet_params(pa_unet h, float* dst, const float* src);
int pa_unet_forward(pa_unet h, const float* x, int batch, float* y);
/* mean-squared-error loss over all output channels and its gradient w.r.t. the
 * parameters (grad may be NULL); loss on the host */
int pa_unet_gradient(pa_unet h, const float* x, const float* target, int batch, float* grad, double* loss);
/* pa_unet_gradient followed by one Adam step (beta 0.9, 0.999, eps 1e-8) at rate lr */
int pa_unet_train_step(pa_unet h, const float* x, const float* target, int batch, float lr, double* loss);

#ifdef __cplusplus
}
#endif
#endif /* PLATO_MI355X_H */
