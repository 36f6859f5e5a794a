// Jacobian assembly on a uniform box lattice (every hex split into the same six tets).
//
// On such a mesh each element is one of six lattice tets, so its stiffness block for a
// vertex pair (a, b) is w_e * B[t][a][b], with B the unit-weight block of lattice tet t
// (vol (lam g_a g_b^T + mu g_b g_a^T + mu (g_a.g_b) I)) and w_e the penalised density
// weight of the element.  A graph block K_rc is then a weighted sum of at most a few
// constant blocks: the elements around the edge (r, c) are fixed lattice offsets from r.
// The kernel needs no connectivity, coordinates or element records; it reads the graph,
// one packed lattice position per node and the element weights, and writes each CRS
// value once.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pa_kernels.hpp"

namespace pa {

constexpr int kBoxMaxRow = 16;     // graph row length limit (16 stencil lanes per row)
constexpr int kBoxRows = 32;       // rows per 256-thread work-group tile (8 lanes per row)
constexpr int kBoxMaxTerms = 6;    // element terms per off-diagonal block
constexpr int kBoxDiagLane = 15;   // stencil lane of the diagonal block
constexpr int kBoxMaxAxis = 1024;  // nodes per axis (10-bit packed lattice positions)

// Per stencil neighbour l (< 15; lane 15 is the diagonal): its lattice offset, the
// number of element terms and, per term e, the element-weight offset from six times the
// row's padded hex index (6 * (-hex offset) + t) and the unit-weight block P of (t, a, b).
struct BoxStencil {
  int count[16];
  int off[16][kBoxMaxTerms];
  int uoff[16];        // padded-lattice offset of the neighbour: d0 + px (d1 + py d2)
  int8_t d[16][3];
};

struct BoxArgs {
  const int* rowmap;
  const int* col;
  const int* gpack;        // per node: i | j << 10 | k << 20 (lattice position)
  const unsigned long long* perm;  // per row: CRS slot of stencil lane l in bits 4l..4l+3
  const uint8_t* corners;  // [6][4] corner of each local vertex of the lattice tets (bit 0 x, 1 y, 2 z)
  const double* z;
  double* W;               // [(nx+2)(ny+2)(nz+2)][6] element weights per padded hex, zero on the ghost layer
  const double* B;         // [16][6][9] material blocks per stencil lane and term at unit weight
  const BoxStencil* S;
  const double* u;
  double* ulat;            // [padded lattice][3] u at node lattice points (padded like the hexes)
  double* zlat;            // [padded lattice] z at node lattice points
  const int* lidx;         // per node: its entry in the load list, or -1
  const double* load_f;    // [3 * loads] nodal loads f
  double* vals;
  double* R;
  int nrows;
  int nx, ny, nz;          // hexes per axis
  int seg_cap;             // k_box_jac_ws: blocks of one staged tile (kBoxRows x longest row)
  Material m;
  // k_box_res_kuhn: mesh tet index of each Kuhn path, 1 / spacing per axis, tet volume
  int tperm[6];
  double ih[3];
  double vol;
};

// z (and u when ulat is set) of every node at its padded lattice position (one thread
// per node: coalesced reads, scattered writes).
__global__ void k_box_scatter(BoxArgs A);

// Element weights of every padded hex from the lattice-ordered z into W (T = double: the
// assembly's weights; T = float: the FP32 matrix-free operator's); ghost hexes get zero
// weights.  One wave per run of 64 padded hexes along x.
// SIMP_INT: the material's penalty is SIMP with an integer exponent (Material::pint > 0).
template <bool SIMP_INT, typename T>
__global__ void k_box_w(BoxArgs A, T* W);

// B[l][e] = lam P + mu (P^T + tr(P) I) from the unit-weight blocks P[l][e] = vol g_a g_b^T,
// and its FP32 copy Bf.
__global__ void k_box_material(const double* P, double* B, float* Bf, double lam, double mu);

// K values of every graph row (and R = K u - f when WITH_R); 32-row tiles per work-group.
template <bool WITH_R>
__global__ void k_box_jac(BoxArgs A);

// The same values with the work-group's waves specialised: four waves compute tile k+1
// into one LDS buffer while two others stream tile k out of the other, so the waves
// that wait on gathers never hold outstanding stores and the waves that store never
// wait on loads.  kBoxWsThreads threads; dynamic LDS of box_jac_ws_lds(seg_cap) bytes.
constexpr int kBoxWsThreads = 384;
template <bool WITH_R>
__global__ void k_box_jac_ws(BoxArgs A);
// k_box_jac_ws with the computing waves' gathers issued one tile ahead (right after the
// previous tile's arithmetic), for lattices whose weights and displacements do not fit
// the last-level cache; R is written without the load (the caller subtracts f).  Same
// threads and LDS.
template <bool WITH_R>
__global__ void k_box_jac_ws_ahead(BoxArgs A);
inline size_t box_jac_ws_lds(int seg_cap) { return (size_t)(2 * 9 * seg_cap + 2 * 3 * kBoxRows) * sizeof(double); }

// FP32 matrix-free operator on the lattice (the mixed-precision PCG's K p and the
// multigrid's level-0 smoother operator): y = K p with the direction update
// p = z + beta p_prev fused, as in the chunk-based FP32 apply (constrained dofs: p = 0,
// y = p).  K is never formed: each row sums, over its stencil neighbours c, the blocks
// K_rc = sum_e w_e B_e (FP32 weights and unit blocks) times p_c - p_r.
struct BoxMfArgs {
  const int* gpack;        // per node: i | j << 10 | k << 20
  const BoxStencil* S;
  const float* W;          // [padded hexes][6] FP32 element weights, zero on the ghost layer
                           // (mesh tet order; Kuhn-path order for k_boxmf_apply_kuhn)
  const float* B;          // [16][6][9] FP32 unit-weight blocks
  const uint8_t* bcmask;   // bit i: dof 3n+i constrained
  const double* z;         // FP64 vectors in node order
  const double* p_prev;
  double* p_cur;           // null: no FP64 copy of p (constrained rows give y = 0)
  int first;               // p = z (else p = z + scal[kBeta] p_prev)
  const double* scal;
  float4* plat;            // [padded lattice] p (FP32) at the node's lattice point
  double* y;
  double* partials;        // one p.y partial per work-group of k_boxmf_apply
  int nrows;
  int nx, ny, nz;          // hexes per axis
  // k_boxmf_apply_kuhn (W in Kuhn-path order): 1 / spacing per axis, tet volume, Lame
  // constants
  float ih[3];
  float vol, lam, mu;
};

// The six tets of a hex split along its main diagonal: tet q visits corners 0,
// e_a, e_a + e_b, e_a + e_b + e_c for the axis path (a, b, c) = kKuhnPath[q].
constexpr int kKuhnPath[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

// R = K u - f (f subtracted when sub_load) without forming K, in FP64, from the element
// weights A.W and u at the lattice points (A.ulat); the rows sum in a fixed order.
__global__ void k_box_res(BoxArgs A, int sub_load);

// The same residual tet by tet on lattices split into the six main-diagonal Kuhn tets
// (see k_boxmf_apply_kuhn), all in FP64: one row per lane, u at the 15 lattice points its
// 24 tets reach (A.ulat zero on the ghost layer and one padded plane and row past the
// end), each tet's force vol w sigma g_m added in a fixed order.
__global__ void k_box_res_kuhn(BoxArgs A, int sub_load);

// Inverted 3x3 diagonal blocks from the same Kuhn tets, in FP64: K_rr = sum over the
// row's 24 tets of vol w ((lambda + mu) g g^T + mu |g|^2 I), g the row vertex's gradient.
__global__ void k_box_diag_inv_kuhn(BoxArgs A, const uint8_t* bcmask, double* dinv);

// Inverted 3x3 diagonal blocks of K(z) (the block-Jacobi preconditioner; constrained
// dofs identity) from the element weights A.W: K_rr = -sum_{c != r} K_rc summed over
// the row's lanes in a fixed order.
__global__ void k_box_diag_inv(BoxArgs A, const uint8_t* bcmask, double* dinv);

// p_cur (FP64, node order) and its FP32 copy at the lattice points (one thread per node).
__global__ void k_boxmf_scatter(BoxMfArgs A);

// y = K p_cur and the work-group's p.y partial; 8 lanes per row, 32-row tiles, a
// persistent XCD-contiguous grid like k_box_jac.
__global__ void k_boxmf_apply(BoxMfArgs A);

// The same product for lattices split into the six main-diagonal Kuhn tets, one row per
// lane from the 24 tets around its lattice point: in tet (a, b, c) column a of the
// displacement gradient is (p(e_a) - p(0)) / h_a, column b (p(e_a + e_b) - p(e_a)) / h_b,
// column c the last edge over h_c; the force on local vertex m is
// vol w sigma g_m with g_m = e_{path[m-1]} / h - e_{path[m]} / h (terms present for
// 0 < m and m < 3).  No block tables; the lattice p must be zero on the ghost layer.
// Persistent XCD-contiguous grid over the rows, one p.y partial per work-group.
__global__ void k_boxmf_apply_kuhn(BoxMfArgs A);

}  // namespace pa
