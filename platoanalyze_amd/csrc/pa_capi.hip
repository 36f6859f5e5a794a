// C ABI of the Tet4 SIMP elastostatics library (declared in include/plato_mi355x.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/plato_mi355x.h"
#include "pa_host.hpp"
#include "pa_box.hpp"
#include "pa_deflate.hpp"
#include "pa_kernels.hpp"
#include "pa_mf.hpp"
#include "pa_mg.hpp"
#include "pa_reduce.hpp"
#include "pa_solver.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define PA_HIP(expr)                                                                 \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) return fail(PA_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// vol * max_a |grad N_a|^2 of a Tet4, rounded up to float (a bound on the magnitude of
// the element's unit-weight stiffness entries divided by |lambda| + 2 |mu|).
float element_grad_bound(const double* X, const int32_t* v) {
  double x[4][3];
  for (int a = 0; a < 4; ++a)
    for (int d = 0; d < 3; ++d) x[a][d] = X[3 * (size_t)v[a] + d];
  const double j11 = x[0][0] - x[3][0], j12 = x[1][0] - x[3][0], j13 = x[2][0] - x[3][0];
  const double j21 = x[0][1] - x[3][1], j22 = x[1][1] - x[3][1], j23 = x[2][1] - x[3][1];
  const double j31 = x[0][2] - x[3][2], j32 = x[1][2] - x[3][2], j33 = x[2][2] - x[3][2];
  const double det = j11 * j22 * j33 + j12 * j23 * j31 + j13 * j21 * j32 - j11 * j23 * j32 - j12 * j21 * j33 -
                     j13 * j22 * j31;
  if (!(det != 0.0) || !std::isfinite(det)) return 0.0f;
  const double inv = 1.0 / det;
  double g[4][3] = {{(j22 * j33 - j23 * j32) * inv, (j13 * j32 - j12 * j33) * inv, (j12 * j23 - j13 * j22) * inv},
                    {(j23 * j31 - j21 * j33) * inv, (j11 * j33 - j13 * j31) * inv, (j13 * j21 - j11 * j23) * inv},
                    {(j21 * j32 - j22 * j31) * inv, (j12 * j31 - j11 * j32) * inv, (j11 * j22 - j12 * j21) * inv},
                    {0, 0, 0}};
  for (int d = 0; d < 3; ++d) g[3][d] = -(g[0][d] + g[1][d] + g[2][d]);
  double gm = 0.0;
  for (int a = 0; a < 4; ++a) gm = std::max(gm, g[a][0] * g[a][0] + g[a][1] * g[a][1] + g[a][2] * g[a][2]);
  const double b = std::fabs(det) / 6.0 * gm * (1.0 + 1e-6);
  return std::nextafter((float)b, std::numeric_limits<float>::infinity());
}

// LDS budget of one assembly work-group (two work-groups per CU fit in 160 KiB).
constexpr int kMaxChunkBlocks = 896;
constexpr int kMaxChunkRows = 128;
constexpr int kMaxChunkNodes = 256;  // node table entries (rows + coupled nodes) per chunk

template <typename T>
struct DevArray {
  T* p = nullptr;
  size_t n = 0;
  ~DevArray() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    release();
    n = count;
    if (count == 0) return hipSuccess;
    return hipMalloc(&p, count * sizeof(T));
  }
  hipError_t upload(const std::vector<T>& v) {
    hipError_t e = alloc(v.size());
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  }
};

}  // namespace

// One level of the geometric multigrid hierarchy of a box mesh.  Level 0 is the
// problem itself; coarser levels own a problem handle on the coarsened box.
// Lattice arrays (element weights and lattice u) above this size take k_box_jac_ws_ahead:
// about the last-level cache (Infinity Cache) of the part.
constexpr size_t kBoxAheadBytes = (size_t)256 << 20;

struct MgLevel {
  pa_handle h = nullptr;  // owned for levels >= 1
  bool owned = false;
  pa::BoxDims dims{0, 0, 0};
  int nn = 0;
  DevArray<int32_t> grid;          // (i, j, k) per node
  DevArray<int32_t> node_of_grid;  // node per grid point
  DevArray<double> r, x, y, d, t, rr, w, vals, dinv;  // w: nodal stiffness weight
  DevArray<float> vals32;          // FP32 copy of vals for the mixed solve's smoothing
  DevArray<double> pw;             // smoothing and spectral-estimate vector
  double inv_theta = 0.0, c1 = 0.0, c2 = 0.0, lam = 0.0, lmax = 0.0;
  ~MgLevel();
};

struct pa_problem {
  int nn = 0, ne = 0;
  std::vector<double> coords;
  std::vector<int32_t> conn;
  double E = 1.0, nu = 0.3;
  double p = 3.0, rmin = 1e-9;
  int pmodel = 0;       // pa::PenaltyModel
  double plen = 1.0;    // Heaviside regularization length
  std::vector<int32_t> bc_dofs;
  std::vector<double> bc_vals;
  std::vector<double> f;  // dense load, filled by pa_add_traction
  std::vector<int32_t> trac_fv;  // traction faces (3 vertices each) and their vectors,
  std::vector<double> trac_t;    // kept for the load's shape derivative
  bool user_graph = false;
  pa::Graph graph;
  bool finalized = false;
  hipStream_t stream = nullptr;

  // device data
  int code_bits = 4;
  pa::Chunks chunks;
  DevArray<double> X;
  DevArray<int32_t> dconn;
  DevArray<uint32_t> codes;
  DevArray<int32_t> rowmap, colidx, diag_slot;
  DevArray<int32_t> chunk_rows, list_offs;
  DevArray<uint32_t> list;
  DevArray<int32_t> chunk_node_offs, chunk_nodes;
  DevArray<uint16_t> ent_ln;
  DevArray<uint32_t> ent_code;
  DevArray<uint16_t> tslot;
  DevArray<int32_t> load_nodes, load_offs, load_row;
  DevArray<double> load_f;
  int nload = 0;
  DevArray<int32_t> bc_d;
  DevArray<double> bc_v;
  DevArray<uint8_t> bcmask;
  double f_free_norm2 = 0.0;
  // mixed-precision solver workspace
  DevArray<float4> m_xz;
  DevArray<double> m_dinv, m_r, m_z, m_p, m_p2, m_q, m_R;
  DevArray<double> partials;  // kNumPartialSlots x kRedBlocks
  DevArray<double> scal;
  DevArray<double> chunk_partials;
  DevArray<float> chunk_gbound;  // per chunk: bound of vol |grad N|^2 (fixed-point Jacobian)
  // deflated solve (pa_solution_deflated): deflation vectors, their images, partials,
  // E = V^T A V, its inverse, mu
  DevArray<double> d_V, d_AV, d_P, d_E, d_Einv, d_mu;
  // solver workspace (lazily allocated)
  DevArray<double> w_r, w_z, w_p, w_q, w_dinv, w_vals, w_adj, w_adjz, w_K;
  // shape derivative of the traction load: value positions in the transposed dR/dX
  // matrix and their (constant) contributions, built on first use
  bool gx_built = false;
  DevArray<int64_t> gx_idx;
  DevArray<double> gx_val;

  // geometric multigrid (box meshes, pa_set_box)
  bool has_box = false;
  pa::BoxDims box{0, 0, 0};       // node counts per axis
  double box_lo[3] = {0, 0, 0}, box_len[3] = {0, 0, 0};
  std::vector<int32_t> box_grid;  // (i, j, k) per node, host
  std::vector<std::unique_ptr<MgLevel>> mg;
  DevArray<double> mg_Ainv;
  // Smoothing (measured at 128x64x64 and 256x128x128): on level 0 one damped block-Jacobi
  // step before and after the coarse correction (Chebyshev degree 1: 52-54 iterations at
  // 3 operator applications each beat degree 2's 48 at 5; unequal degrees break the
  // symmetry PCG needs: 72-74), with the FP32 matrix-free operator also under the FP64
  // solve (its CG operator stays the FP64 CRS matrix); degree 2 on coarser levels
  // (degree 1 there: 74-76 iterations); interval [lambda / 60, 1.1 lambda].
  bool mg_fp32_smooth = true;
  int mg_pre0 = 1, mg_post0 = 1, mg_degc = 2;
  bool mg_f32 = true;  // mixed solve: coarse smoothing on FP32 matrix copies (PA_MG_F32=0: FP64)
  double mg_ratio = 60.0;
  DevArray<double> g_r, g_z, g_p, g_q, g_pc;  // PCG vectors of the multigrid solve
  // lattice assembly (pa_box.hip): enabled by box_fast_setup when the box's hexes all
  // carry the same six lattice tets, the nodes sit on the uniform lattice and the graph
  // rows fit 16 blocks inside the 27-point stencil
  bool box_fast = false;
  DevArray<int32_t> box_gpack;
  DevArray<uint8_t> box_corners;
  DevArray<int32_t> box_lidx;
  DevArray<unsigned long long> box_perm;
  DevArray<double> box_P, box_B, box_W, box_ulat, box_zlat;
  DevArray<pa::BoxStencil> box_S;
  double box_lam = 0.0, box_mu = 0.0;  // material of box_B (rebuilt when it changes)
  int box_grid_k = 1024, box_grid_r = 768;  // persistent grids of the lattice kernel (K, K and R)
  // FP32 matrix-free operator on the lattice (k_boxmf_apply): FP32 unit blocks, per-solve
  // FP32 element weights, p at the lattice points, one p.y partial per work-group
  DevArray<float> box_Bf, box_W32;
  DevArray<float4> box_plat;
  DevArray<double> box_mf_part;
  int box_grid_mf = 1024;
  bool box_mf = true;  // PA_BOX_MF=0 keeps the chunk-based FP32 apply on lattice meshes
  // k_boxmf_apply_kuhn in place of k_boxmf_apply: hexes split into the six main-diagonal
  // Kuhn tets (PA_BOX_MF_KUHN=0 keeps k_boxmf_apply); its FP32 weights are stored in
  // Kuhn-path order (box_corners_kuhn: the corner sets in kKuhnPath order)
  bool box_kuhn = false;
  DevArray<uint8_t> box_corners_kuhn;
  // k_box_res_kuhn and k_box_diag_inv_kuhn in place of k_box_res and k_box_diag_inv on
  // the same lattices (PA_BOX_RES_KUHN=0 keeps the stencil kernels); mesh tet index of
  // each Kuhn path
  bool box_res_kuhn = false;
  int box_tperm[6] = {0, 0, 0, 0, 0, 0};
  double box_step[3] = {0.0, 0.0, 0.0};
  // k_box_jac_ws (store and compute waves specialised; PA_BOX_WS=0 selects k_box_jac)
  bool box_ws = true;
  int box_seg_cap = 0, box_grid_ws = 512;
  // k_box_jac_ws_ahead in place of k_box_jac_ws: lattices whose element weights and
  // lattice u exceed kBoxAheadBytes (PA_BOX_AHEAD=0/1 overrides)
  bool box_ahead = false;
  bool box_res = true;  // lattice FP64 residual (k_box_res); PA_BOX_RES=0 keeps the chunk kernel
  int box_grid_res = 1024;
  double* host_scal = nullptr;

  pa::Material material() const {
    const double c = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
    pa::Material m;
    m.c00 = c * (1.0 - nu);
    m.c01 = c * nu;
    m.c33 = 1.0 / 2.0 * c * (1.0 - 2.0 * nu);
    m.p = p;
    m.rmin = rmin;
    m.pint = (p >= 1.0 && p <= 8.0 && p == std::floor(p)) ? (int)p : 0;
    m.model = pmodel;
    m.len = plen;
    return m;
  }
  pa::ChunkArgs chunk_args() const {
    pa::ChunkArgs a{};
    a.conn = reinterpret_cast<const int4*>(dconn.p);
    a.codes = codes.p;
    a.X = X.p;
    a.rowmap = rowmap.p;
    a.colidx = colidx.p;
    a.chunk_rows = chunk_rows.p;
    a.list_offs = list_offs.p;
    a.list = list.p;
    a.node_offs = chunk_node_offs.p;
    a.nodes = chunk_nodes.p;
    a.ent_ln = ent_ln.p;
    a.ent_code = ent_code.p;
    a.tslot = tslot.p;
    a.max_nodes = chunks.max_nodes;
    a.partials = chunk_partials.p;
    a.gbound = chunk_gbound.p;
    a.load_offs = load_offs.p;
    a.load_row = load_row.p;
    a.load_f = load_f.p;
    a.sub_load = 0;
    a.nchunks = chunks.count;
    a.max_rows = chunks.max_rows;
    a.max_blocks = chunks.max_blocks;
    a.m = material();
    return a;
  }
  ~pa_problem() {
    if (host_scal) (void)hipHostFree(host_scal);
  }
};

MgLevel::~MgLevel() {
  if (owned) delete h;
}

static int box_fast_setup(pa_handle h);
static pa::BoxArgs box_args(pa_handle h, const double* u, const double* z, double* vals, double* R);
template <typename T>
static void box_weights(pa_handle h, const pa::BoxArgs& a, T* W);
static int assemble_jacobian(pa_handle h, const double* u, const double* z, double* vals, double* R);

static int grid_for(size_t n, int per_block = 256, int cap = 4096) {
  size_t g = (n + per_block - 1) / per_block;
  return (int)std::max<size_t>(1, std::min<size_t>(g, (size_t)cap));
}

static int ready(pa_handle h) {
  if (!h) return fail(PA_ERR_ARG, "null handle");
  if (!h->finalized) return fail(PA_ERR_STATE, "pa_finalize has not been called");
  return PA_OK;
}

extern "C" {

const char* pa_last_error(void) { return g_err.c_str(); }

int pa_build_box(int nx, int ny, int nz, double lx, double ly, double lz, int* num_nodes, int* num_elems,
                 double* coords, int32_t* conn) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || !(lx > 0) || !(ly > 0) || !(lz > 0))
    return fail(PA_ERR_ARG, "box dimensions must be positive");
  const int64_t nv = (int64_t)(nx + 1) * (ny + 1) * (nz + 1), nel = 6LL * nx * ny * nz;
  if (nel >= (1LL << 28)) return fail(PA_ERR_UNSUPPORTED, "more than 2^28 elements");
  if (num_nodes) *num_nodes = (int)nv;
  if (num_elems) *num_elems = (int)nel;
  if (!coords && !conn) return PA_OK;
  if (!coords || !conn) return fail(PA_ERR_ARG, "coords and conn must both be given");
  pa::BoxMesh m = pa::build_box(nx, ny, nz, lx, ly, lz);
  std::memcpy(coords, m.coords.data(), m.coords.size() * sizeof(double));
  std::memcpy(conn, m.conn.data(), m.conn.size() * sizeof(int32_t));
  return PA_OK;
}

int pa_nodes_on_plane(int num_nodes, const double* coords, int axis, double value, double tol, int32_t* out,
                      int* count) {
  if (!coords || !count || axis < 0 || axis > 2) return fail(PA_ERR_ARG, "bad arguments");
  auto v = pa::nodes_on_plane(coords, num_nodes, axis, value, tol);
  *count = (int)v.size();
  if (out) std::memcpy(out, v.data(), v.size() * sizeof(int32_t));
  return PA_OK;
}

int pa_faces_on_plane(int num_elems, const int32_t* conn, const double* coords, int axis, double value,
                      double tol, int32_t* face_verts, int32_t* face_elems, int* count) {
  if (!conn || !coords || !count || axis < 0 || axis > 2) return fail(PA_ERR_ARG, "bad arguments");
  std::vector<int32_t> fv, fe;
  pa::faces_on_plane(conn, num_elems, coords, axis, value, tol, fv, fe);
  *count = (int)fe.size();
  if (face_verts) std::memcpy(face_verts, fv.data(), fv.size() * sizeof(int32_t));
  if (face_elems) std::memcpy(face_elems, fe.data(), fe.size() * sizeof(int32_t));
  return PA_OK;
}

int pa_node_graph(int nn, int ne, const int32_t* conn, int32_t* rowmap, int32_t* colidx, int* nnz) {
  if (nn <= 0 || ne <= 0 || !conn || !rowmap || !nnz) return fail(PA_ERR_ARG, "pa_node_graph: bad arguments");
  for (int64_t k = 0; k < 4LL * ne; ++k)
    if (conn[k] < 0 || conn[k] >= nn) return fail(PA_ERR_MESH, "connectivity index out of range");
  pa::Graph g = pa::build_graph(conn, ne, nn);
  std::memcpy(rowmap, g.rowmap.data(), g.rowmap.size() * sizeof(int32_t));
  *nnz = g.rowmap[nn];
  if (colidx) std::memcpy(colidx, g.colidx.data(), g.colidx.size() * sizeof(int32_t));
  return PA_OK;
}

int pa_create(pa_handle* h, int num_nodes, int num_elems, const double* coords, const int32_t* conn) {
  if (!h || !coords || !conn || num_nodes <= 0 || num_elems <= 0)
    return fail(PA_ERR_ARG, "pa_create: null pointer or empty mesh");
  if (num_elems >= (1 << 28)) return fail(PA_ERR_UNSUPPORTED, "more than 2^28 elements");
  for (int64_t k = 0; k < 4LL * num_elems; ++k)
    if (conn[k] < 0 || conn[k] >= num_nodes) return fail(PA_ERR_MESH, "connectivity index out of range");
  auto p = std::make_unique<pa_problem>();
  p->nn = num_nodes;
  p->ne = num_elems;
  p->coords.assign(coords, coords + 3LL * num_nodes);
  p->conn.assign(conn, conn + 4LL * num_elems);
  p->f.assign(3LL * num_nodes, 0.0);
  *h = p.release();
  return PA_OK;
}

int pa_destroy(pa_handle h) {
  delete h;
  return PA_OK;
}

int pa_set_material(pa_handle h, double E, double nu) {
  if (!h) return fail(PA_ERR_ARG, "null handle");
  if (!(E > 0) || !(nu > -1.0 && nu < 0.5)) return fail(PA_ERR_ARG, "material needs E>0 and -1<nu<0.5");
  h->E = E;
  h->nu = nu;
  return PA_OK;
}

int pa_set_penalty(pa_handle h, double exponent, double minimum_value) {
  return pa_set_penalty_model(h, pa::kPenaltySimp, exponent, minimum_value, 1.0);
}

int pa_set_penalty_model(pa_handle h, int model, double exponent, double minimum_value, double length) {
  if (!h) return fail(PA_ERR_ARG, "null handle");
  if (!std::isfinite(exponent) || !std::isfinite(minimum_value) || !std::isfinite(length))
    return fail(PA_ERR_ARG, "non-finite penalty");
  if (model < pa::kPenaltySimp || model > pa::kPenaltyNone) return fail(PA_ERR_ARG, "unknown penalty model");
  if (model == pa::kPenaltyHeaviside && !(length > 0.0)) return fail(PA_ERR_ARG, "Heaviside length must be > 0");
  h->pmodel = model;
  h->p = exponent;
  h->rmin = minimum_value;
  h->plen = length;
  return PA_OK;
}

int pa_add_essential_bc(pa_handle h, int num_nodes, const int32_t* nodes, int dof, double value) {
  if (!h) return fail(PA_ERR_ARG, "null handle");
  if (h->finalized) return fail(PA_ERR_STATE, "boundary conditions must be added before pa_finalize");
  if (num_nodes < 0 || (num_nodes > 0 && !nodes) || dof < 0 || dof > 2)
    return fail(PA_ERR_ARG, "essential BC: bad node list or dof");
  for (int k = 0; k < num_nodes; ++k) {
    if (nodes[k] < 0 || nodes[k] >= h->nn) return fail(PA_ERR_ARG, "essential BC node out of range");
    h->bc_dofs.push_back(3 * nodes[k] + dof);
    h->bc_vals.push_back(value);
  }
  return PA_OK;
}

int pa_add_traction(pa_handle h, int num_faces, const int32_t* fv, const int32_t* fe, const double t[3]) {
  if (!h) return fail(PA_ERR_ARG, "null handle");
  if (h->finalized) return fail(PA_ERR_STATE, "loads must be added before pa_finalize");
  if (num_faces < 0 || (num_faces > 0 && !fv) || !t) return fail(PA_ERR_ARG, "traction: bad arguments");
  for (int k = 0; k < 3; ++k)
    if (!std::isfinite(t[k])) return fail(PA_ERR_ARG, "traction: non-finite value");
  // SurfaceLoadIntegral: weight = cub * |(x0-x2) x (x1-x2)| with the tet rule
  // weight 1/6, i.e. area/3 per face vertex; R gets -weight*t, so f gets +weight*t.
  const double cub = 1.0 / 2.0 / 3.0;
  for (int k = 0; k < num_faces; ++k) {
    int v[3];
    for (int a = 0; a < 3; ++a) {
      v[a] = fv[3 * k + a];
      if (v[a] < 0 || v[a] >= h->nn) return fail(PA_ERR_ARG, "traction face vertex out of range");
    }
    if (fe && (fe[k] < 0 || fe[k] >= h->ne)) return fail(PA_ERR_ARG, "traction face element out of range");
    const double* c = h->coords.data();
    double j0[3], j1[3];
    for (int d = 0; d < 3; ++d) {
      j0[d] = c[3 * v[0] + d] - c[3 * v[2] + d];
      j1[d] = c[3 * v[1] + d] - c[3 * v[2] + d];
    }
    const double j23 = j0[1] * j1[2] - j0[2] * j1[1];
    const double j31 = j0[2] * j1[0] - j0[0] * j1[2];
    const double j12 = j0[0] * j1[1] - j0[1] * j1[0];
    const double w = cub * std::sqrt(j23 * j23 + j31 * j31 + j12 * j12);
    for (int a = 0; a < 3; ++a)
      for (int d = 0; d < 3; ++d) h->f[3 * (size_t)v[a] + d] += w * t[d];
    h->trac_fv.insert(h->trac_fv.end(), v, v + 3);
    h->trac_t.insert(h->trac_t.end(), t, t + 3);
  }
  return PA_OK;
}

int pa_set_graph(pa_handle h, const int32_t* rm, const int32_t* ci) {
  if (!h || !rm || !ci) return fail(PA_ERR_ARG, "pa_set_graph: null pointer");
  if (h->finalized) return fail(PA_ERR_STATE, "graph must be set before pa_finalize");
  h->graph.rowmap.assign(rm, rm + h->nn + 1);
  if (h->graph.rowmap[h->nn] < 0) return fail(PA_ERR_MESH, "graph rowmap is negative");
  h->graph.colidx.assign(ci, ci + h->graph.rowmap[h->nn]);
  std::string err = pa::check_graph(h->graph, h->conn.data(), h->ne, h->nn);
  if (!err.empty()) return fail(PA_ERR_MESH, err);
  h->user_graph = true;
  return PA_OK;
}

int pa_finalize(pa_handle h) {
  if (!h) return fail(PA_ERR_ARG, "null handle");
  if (h->finalized) return PA_OK;
  if (!h->user_graph) h->graph = pa::build_graph(h->conn.data(), h->ne, h->nn);
  const pa::Graph& g = h->graph;
  if (g.max_row_len <= 16) h->code_bits = 4;
  else if (g.max_row_len <= 256) h->code_bits = 8;
  else return fail(PA_ERR_UNSUPPORTED, "a graph row has more than 256 blocks");
  std::string err;
  std::vector<uint32_t> codes = pa::slot_codes(g, h->conn.data(), h->ne, h->code_bits, err);
  if (!err.empty()) return fail(PA_ERR_MESH, err);
  pa::NodeElems adj = pa::node_elems(h->conn.data(), h->ne, h->nn);
  int max_blocks = kMaxChunkBlocks;
  if (const char* env = std::getenv("PA_CHUNK_BLOCKS")) max_blocks = std::max(64, std::min(4096, std::atoi(env)));
  int max_rows = kMaxChunkRows, max_nodes = kMaxChunkNodes;  // tuning knobs, bounded by LDS below
  if (const char* env = std::getenv("PA_CHUNK_ROWS")) max_rows = std::max(16, std::min(1024, std::atoi(env)));
  if (const char* env = std::getenv("PA_CHUNK_NODES")) max_nodes = std::max(64, std::min(4096, std::atoi(env)));
  const int wpc = h->code_bits == 4 ? 2 : 4;
  h->chunks = pa::build_chunks(g, adj, h->conn.data(), codes, wpc, max_blocks, max_rows, max_nodes);
  for (int mode = 0; mode < pa::kNumChunkModes; ++mode)
    if (pa::chunk_lds_bytes(mode, h->chunks.max_rows, h->chunks.max_blocks, h->chunks.max_nodes) > 160 * 1024)
      return fail(PA_ERR_UNSUPPORTED, "chunk limits exceed the 160 KiB of LDS of a work-group");
  if (h->chunks.max_blocks > max_blocks || h->chunks.max_nodes > max_nodes)
    return fail(PA_ERR_UNSUPPORTED, "a single graph row exceeds the LDS chunk budget");
  std::vector<int32_t> diag(h->nn);
  for (int r = 0; r < h->nn; ++r) {
    diag[r] = -1;
    for (int k = g.rowmap[r]; k < g.rowmap[r + 1]; ++k)
      if (g.colidx[k] == r) diag[r] = k;
    if (diag[r] < 0) return fail(PA_ERR_MESH, "graph row without diagonal");
  }
  // loaded nodes in row order, grouped by chunk (chunks are consecutive row ranges)
  std::vector<int32_t> ln, loffs(h->chunks.count + 1, 0), lrow;
  std::vector<double> lf;
  for (int c = 0; c < h->chunks.count; ++c) {
    for (int n = h->chunks.row_begin[c]; n < h->chunks.row_begin[c + 1]; ++n) {
      const double* fn = &h->f[3 * (size_t)n];
      if (fn[0] != 0.0 || fn[1] != 0.0 || fn[2] != 0.0) {
        ln.push_back(n);
        lrow.push_back(n - h->chunks.row_begin[c]);
        lf.insert(lf.end(), fn, fn + 3);
      }
    }
    loffs[c + 1] = (int32_t)ln.size();
  }
  h->nload = (int)ln.size();
  PA_HIP(h->X.upload(h->coords));
  PA_HIP(h->dconn.upload(h->conn));
  PA_HIP(h->codes.upload(codes));
  PA_HIP(h->rowmap.upload(g.rowmap));
  PA_HIP(h->colidx.upload(g.colidx));
  PA_HIP(h->diag_slot.upload(diag));
  PA_HIP(h->chunk_rows.upload(h->chunks.row_begin));
  PA_HIP(h->list_offs.upload(h->chunks.list_offs));
  PA_HIP(h->list.upload(h->chunks.list));
  PA_HIP(h->chunk_node_offs.upload(h->chunks.node_offs));
  PA_HIP(h->chunk_nodes.upload(h->chunks.nodes));
  PA_HIP(h->ent_ln.upload(h->chunks.ent_ln));
  PA_HIP(h->ent_code.upload(h->chunks.ent_code));
  PA_HIP(h->tslot.upload(h->chunks.tslot));
  PA_HIP(h->load_nodes.upload(ln));
  PA_HIP(h->load_offs.upload(loffs));
  PA_HIP(h->load_row.upload(lrow));
  PA_HIP(h->load_f.upload(lf));
  PA_HIP(h->bc_d.upload(h->bc_dofs));
  PA_HIP(h->bc_v.upload(h->bc_vals));
  {
    std::vector<uint8_t> bm(h->nn, 0);
    for (int32_t d : h->bc_dofs) bm[d / 3] |= (uint8_t)(1u << (d % 3));
    double s2 = 0.0;
    for (int n = 0; n < h->nn; ++n)
      for (int i = 0; i < 3; ++i)
        if (!((bm[n] >> i) & 1u)) s2 += h->f[3 * (size_t)n + i] * h->f[3 * (size_t)n + i];
    h->f_free_norm2 = s2;
    PA_HIP(h->bcmask.upload(bm));
  }
  PA_HIP(h->partials.alloc((size_t)pa::kNumPartialSlots * pa::kRedBlocks));
  PA_HIP(h->scal.alloc(pa::kNumScal));
  PA_HIP(hipMemset(h->scal.p, 0, pa::kNumScal * sizeof(double)));
  PA_HIP(h->chunk_partials.alloc(std::max(1, h->chunks.count)));
  {  // per chunk: a bound on vol |grad N|^2 over its elements (fixed-point Jacobian scale)
    std::vector<float> eb(h->ne), gb(std::max(1, h->chunks.count), 0.0f);
    pa::parallel_for(h->ne, [&](int64_t e0, int64_t e1) {
      for (int64_t e = e0; e < e1; ++e) eb[e] = element_grad_bound(h->coords.data(), &h->conn[4 * (size_t)e]);
    });
    for (int c = 0; c < h->chunks.count; ++c)
      for (int k = h->chunks.list_offs[c]; k < h->chunks.list_offs[c + 1]; ++k)
        gb[c] = std::max(gb[c], eb[h->chunks.list[k] & 0x0FFFFFFFu]);
    PA_HIP(h->chunk_gbound.upload(gb));
  }
  PA_HIP(hipHostMalloc(&h->host_scal, pa::kNumScal * sizeof(double), hipHostMallocDefault));
  h->finalized = true;
  if (h->has_box)
    if (int rc = box_fast_setup(h)) return rc;
  return PA_OK;
}

int pa_set_stream(pa_handle h, void* s) {
  if (!h) return fail(PA_ERR_ARG, "null handle");
  h->stream = reinterpret_cast<hipStream_t>(s);
  return PA_OK;
}

int pa_graph_size(pa_handle h, int* num_rows, int* nnz_blocks) {
  if (int rc = ready(h)) return rc;
  if (num_rows) *num_rows = h->graph.num_rows;
  if (nnz_blocks) *nnz_blocks = h->graph.rowmap[h->graph.num_rows];
  return PA_OK;
}

int pa_copy_graph(pa_handle h, int32_t* rm, int32_t* ci, int on_device) {
  if (int rc = ready(h)) return rc;
  const auto& g = h->graph;
  hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  if (rm) PA_HIP(hipMemcpyAsync(rm, h->rowmap.p, (g.num_rows + 1) * sizeof(int32_t), kind, h->stream));
  if (ci) PA_HIP(hipMemcpyAsync(ci, h->colidx.p, g.colidx.size() * sizeof(int32_t), kind, h->stream));
  PA_HIP(hipStreamSynchronize(h->stream));
  return PA_OK;
}

// --------------------------------------------------------------- VectorFunction
static int run_chunk(pa_handle h, int mode, const double* u, const double* z, double* vals, double* vec,
                     bool sub_load = false) {
  pa::ChunkArgs a = h->chunk_args();
  a.sub_load = sub_load ? 1 : 0;
  a.u = u;
  a.z = z;
  a.out_vals = vals;
  a.out_vec = vec;
  PA_HIP(pa::launch_chunk(h->code_bits, mode, a, h->stream));
  return PA_OK;
}

// R = K(z) u - f (f subtracted when sub): on box meshes that take the lattice assembly
// the lattice FP64 residual (element weights of z, u at the lattice points, then the
// rows in a fixed order), else the chunk kernel.
static int residual(pa_handle h, const double* u, const double* z, double* R, bool sub) {
  if (h->box_fast && h->box_res) {
    const pa::BoxArgs a = box_args(h, u, z, nullptr, R);
    box_weights(h, a, a.W);
    if (h->box_res_kuhn) {
      const int G = std::min(4096, ((h->nn + 255) / 256 + pa::kNumXcd - 1) / pa::kNumXcd * pa::kNumXcd);
      pa::k_box_res_kuhn<<<G, 256, 0, h->stream>>>(a, sub ? 1 : 0);
    } else {
      const int nblk = (h->nn + pa::kBoxRows - 1) / pa::kBoxRows;
      const int G = std::min((nblk + pa::kNumXcd - 1) / pa::kNumXcd * pa::kNumXcd, h->box_grid_res);
      pa::k_box_res<<<G, 256, 0, h->stream>>>(a, sub ? 1 : 0);
    }
    PA_HIP(hipGetLastError());
    return PA_OK;
  }
  return run_chunk(h, pa::kModeResidual, u, z, nullptr, R, sub);
}

static int sub_load(pa_handle h, double* R, double s) {
  if (h->nload == 0) return PA_OK;
  pa::k_add_load<<<grid_for(h->nload), 256, 0, h->stream>>>(R, h->load_nodes.p, h->load_f.p, h->nload, s);
  PA_HIP(hipGetLastError());
  return PA_OK;
}

int pa_vf_value(pa_handle h, const double* u, const double* z, double* R) {
  if (int rc = ready(h)) return rc;
  if (!u || !z || !R) return fail(PA_ERR_ARG, "pa_vf_value: null pointer");
  return residual(h, u, z, R, true);
}

int pa_vf_gradient_u(pa_handle h, const double* u, const double* z, double* vals, double* R) {
  if (int rc = ready(h)) return rc;
  if (!z || !vals || (R && !u)) return fail(PA_ERR_ARG, "pa_vf_gradient_u: null pointer");
  return assemble_jacobian(h, u, z, vals, R);
}

int pa_vf_gradient_z(pa_handle h, const double* u, const double* z, double* vals_t) {
  if (int rc = ready(h)) return rc;
  if (!u || !z || !vals_t) return fail(PA_ERR_ARG, "pa_vf_gradient_z: null pointer");
  return run_chunk(h, pa::kModeGradZT, u, z, vals_t, nullptr);
}

// Traction part of dR/dX: R = K u - f with f_a = (|c|/6) t on each face vertex a,
// c = (x0 - x2) x (x1 - x2); d|c|/dx0 = j1 x c^, d|c|/dx1 = c^ x j0, d|c|/dx2 = -(both).
// Contributions to block (config vertex b, residual vertex a), entry (k, i), summed per
// position on the host (they depend on the coordinates and loads only).
static int build_gx_traction(pa_handle h) {
  if (h->gx_built) return PA_OK;
  const pa::Graph& g = h->graph;
  std::vector<std::pair<int64_t, double>> pv;
  const size_t nf = h->trac_fv.size() / 3;
  const double* c = h->coords.data();
  for (size_t f = 0; f < nf; ++f) {
    const int32_t* v = &h->trac_fv[3 * f];
    const double* t = &h->trac_t[3 * f];
    double j0[3], j1[3];
    for (int d = 0; d < 3; ++d) {
      j0[d] = c[3 * (size_t)v[0] + d] - c[3 * (size_t)v[2] + d];
      j1[d] = c[3 * (size_t)v[1] + d] - c[3 * (size_t)v[2] + d];
    }
    const double cx[3] = {j0[1] * j1[2] - j0[2] * j1[1], j0[2] * j1[0] - j0[0] * j1[2], j0[0] * j1[1] - j0[1] * j1[0]};
    const double nrm = std::sqrt(cx[0] * cx[0] + cx[1] * cx[1] + cx[2] * cx[2]);
    if (!(nrm > 0.0)) continue;
    const double n[3] = {cx[0] / nrm, cx[1] / nrm, cx[2] / nrm};
    double dA[3][3];  // d|c| / dx_b
    dA[0][0] = j1[1] * n[2] - j1[2] * n[1];
    dA[0][1] = j1[2] * n[0] - j1[0] * n[2];
    dA[0][2] = j1[0] * n[1] - j1[1] * n[0];
    dA[1][0] = n[1] * j0[2] - n[2] * j0[1];
    dA[1][1] = n[2] * j0[0] - n[0] * j0[2];
    dA[1][2] = n[0] * j0[1] - n[1] * j0[0];
    for (int k = 0; k < 3; ++k) dA[2][k] = -(dA[0][k] + dA[1][k]);
    const double cub = 1.0 / 2.0 / 3.0;
    for (int b = 0; b < 3; ++b)
      for (int a = 0; a < 3; ++a) {
        int slot = -1;
        for (int q = g.rowmap[v[b]]; q < g.rowmap[v[b] + 1]; ++q)
          if (g.colidx[q] == v[a]) slot = q;
        if (slot < 0) return fail(PA_ERR_MESH, "traction face edge missing from the graph");
        for (int k = 0; k < 3; ++k)
          for (int i = 0; i < 3; ++i) pv.emplace_back(9 * (int64_t)slot + 3 * k + i, -cub * dA[b][k] * t[i]);
      }
  }
  std::sort(pv.begin(), pv.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
  std::vector<int64_t> idx;
  std::vector<double> val;
  for (size_t k = 0; k < pv.size(); ++k) {
    if (!idx.empty() && idx.back() == pv[k].first) val.back() += pv[k].second;
    else {
      idx.push_back(pv[k].first);
      val.push_back(pv[k].second);
    }
  }
  PA_HIP(h->gx_idx.upload(idx));
  PA_HIP(h->gx_val.upload(val));
  h->gx_built = true;
  return PA_OK;
}

int pa_vf_gradient_x(pa_handle h, const double* u, const double* z, double* vals_t) {
  if (int rc = ready(h)) return rc;
  if (!u || !z || !vals_t) return fail(PA_ERR_ARG, "pa_vf_gradient_x: null pointer");
  if (int rc = build_gx_traction(h)) return rc;
  if (int rc = run_chunk(h, pa::kModeGradX, u, z, vals_t, nullptr)) return rc;
  const size_t n = h->gx_idx.n;
  if (n) {
    pa::k_scatter_add<<<grid_for(n), 256, 0, h->stream>>>(vals_t, h->gx_idx.p, h->gx_val.p, n);
    PA_HIP(hipGetLastError());
  }
  return PA_OK;
}

int pa_sf_gradient_x(pa_handle h, const double* u, const double* z, double* g) {
  if (int rc = ready(h)) return rc;
  if (!u || !z || !g) return fail(PA_ERR_ARG, "pa_sf_gradient_x: null pointer");
  return run_chunk(h, pa::kModeEnergyGradX, u, z, nullptr, g);
}

static int energy_value(pa_handle h, double* value);
static int ensure_solver_ws(pa_handle h);

// Plato::Geometric::Volume (src/geometric/Volume.hpp:49-96) with its own penalty
// function, through GeometryScalarFunction value / gradient_z / gradient_x
// (src/geometric/GeometryScalarFunction.cpp).
int pa_volume(pa_handle h, const double* z, int model, double exponent, double minimum_value, double length,
              double* value, double* dvdz, double* dvdx) {
  if (int rc = ready(h)) return rc;
  if (!z) return fail(PA_ERR_ARG, "pa_volume: null density");
  if (model < pa::kPenaltySimp || model > pa::kPenaltyNone) return fail(PA_ERR_ARG, "unknown penalty model");
  if (model == pa::kPenaltyHeaviside && !(length > 0.0)) return fail(PA_ERR_ARG, "Heaviside length must be > 0");
  if (int rc = ensure_solver_ws(h)) return rc;
  pa::ChunkArgs a = h->chunk_args();
  a.u = h->w_r.p;  // staged but not used by the volume mode
  a.z = z;
  a.m.model = model;
  a.m.p = exponent;
  a.m.rmin = minimum_value;
  a.m.len = length;
  a.m.pint = (exponent >= 1.0 && exponent <= 8.0 && exponent == std::floor(exponent)) ? (int)exponent : 0;
  a.out_vec = dvdx ? dvdx : h->w_p.p;   // w_p holds 3N doubles
  a.out_vec2 = dvdz ? dvdz : h->w_z.p;  // w_z holds >= N doubles
  PA_HIP(pa::launch_chunk(h->code_bits, pa::kModeVolume, a, h->stream));
  if (value) return energy_value(h, value);
  return PA_OK;
}

// --------------------------------------------------------------- scalar function
static int energy_value(pa_handle h, double* value) {
  pa::k_sum<<<1, 256, 0, h->stream>>>(h->chunk_partials.p, h->chunks.count, h->scal.p, pa::kEnergy);
  PA_HIP(hipGetLastError());
  PA_HIP(hipMemcpyAsync(h->host_scal, h->scal.p, pa::kNumScal * sizeof(double), hipMemcpyDeviceToHost,
                        h->stream));
  PA_HIP(hipStreamSynchronize(h->stream));
  *value = h->host_scal[pa::kEnergy];
  return PA_OK;
}

int pa_sf_value(pa_handle h, const double* u, const double* z, double* value) {
  if (int rc = ready(h)) return rc;
  if (!u || !z || !value) return fail(PA_ERR_ARG, "pa_sf_value: null pointer");
  if (int rc = run_chunk(h, pa::kModeEnergy, u, z, nullptr, nullptr)) return rc;
  return energy_value(h, value);
}

int pa_sf_gradient_u(pa_handle h, const double* u, const double* z, double* g) {
  if (int rc = ready(h)) return rc;
  if (!u || !z || !g) return fail(PA_ERR_ARG, "pa_sf_gradient_u: null pointer");
  return residual(h, u, z, g, false);
}

int pa_sf_gradient_z(pa_handle h, const double* u, const double* z, double* g, double* value) {
  if (int rc = ready(h)) return rc;
  if (!u || !z || !g) return fail(PA_ERR_ARG, "pa_sf_gradient_z: null pointer");
  if (int rc = run_chunk(h, pa::kModeEnergyGradZ, u, z, nullptr, g)) return rc;
  if (value) return energy_value(h, value);
  return PA_OK;
}

// --------------------------------------------------------------- constraints / solve
int pa_apply_constraints(pa_handle h, double* vals, double* rhs, double scale) {
  if (int rc = ready(h)) return rc;
  if (!vals || !rhs) return fail(PA_ERR_ARG, "pa_apply_constraints: null pointer");
  const int nbc = (int)h->bc_dofs.size();
  if (nbc == 0) return PA_OK;
  const int g = grid_for(nbc);
  pa::k_bc_lift<<<g, 256, 0, h->stream>>>(vals, h->rowmap.p, h->colidx.p, h->bc_d.p, h->bc_v.p, nbc, scale, rhs);
  pa::k_bc_zero<<<g, 256, 0, h->stream>>>(vals, h->rowmap.p, h->colidx.p, h->bc_d.p, nbc);
  pa::k_bc_set<<<g, 256, 0, h->stream>>>(rhs, h->bc_d.p, h->bc_v.p, nbc, scale);
  PA_HIP(hipGetLastError());
  return PA_OK;
}

static int ensure_solver_ws(pa_handle h) {
  const size_t ndof = 3 * (size_t)h->nn;
  if (h->w_r.n != ndof) {
    PA_HIP(h->w_r.alloc(ndof));
    PA_HIP(h->w_z.alloc(ndof));
    PA_HIP(h->w_p.alloc(ndof));
    PA_HIP(h->w_q.alloc(ndof));
    PA_HIP(h->w_dinv.alloc(9 * (size_t)h->nn));
  }
  return PA_OK;
}

static int read_scal(pa_handle h, const double* scal) {
  PA_HIP(hipMemcpyAsync(h->host_scal, scal, pa::kNumScal * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  PA_HIP(hipStreamSynchronize(h->stream));
  return PA_OK;
}
static int read_scal(pa_handle h) { return read_scal(h, h->scal.p); }

// Device workspace of the block-Jacobi PCG: four node vectors (3N doubles each), the 3x3
// block-diagonal inverse (9N), kNumPartialSlots x kRedBlocks reduction partials and
// kNumScal scalars.  Every entry is written before it is read.
struct PcgWork {
  double *r, *z, *p, *q, *dinv, *partials, *scal;
};

static int solve_opts(const pa_solver_opts* o, pa_solver_opts& opts, const char* who) {
  opts = o ? *o : pa_solver_opts{1e-10, 100000, 16, 0, 0, 0.1};
  if (opts.check_every <= 0) opts.check_every = 16;
  if (opts.preconditioner != 0)
    return fail(PA_ERR_ARG, std::string(who) + ": only block Jacobi (multigrid needs the density: pa_solution)");
  return PA_OK;
}

// Deflation space of pcg_block_jacobi: k vectors V (zero at the constrained dofs), A V,
// the inverse of E = V^T A V (k x k, device; zero rows and columns for dropped vectors),
// mu (k) and the k x kRedBlocks partials of the projections.
struct Deflation {
  const double* V;
  const double* AV;
  int k;
  const double* Einv;
  double* mu;
  double* P;
};

// y += sign * V E^-1 (X^T v) through the device kernels (X = V or A V)
static int deflate_apply(pa_handle h, const Deflation& D, const double* X, const double* v, double* y, const double* Y,
                         double sign, int G) {
  const size_t ndof = 3 * (size_t)h->nn;
  hipStream_t s = h->stream;
  pa::k_multi_dot<<<G, 256, 0, s>>>(X, D.k, ndof, v, h->nn, D.P);
  pa::k_defl_mu<<<1, 256, 0, s>>>(D.P, G, D.k, D.Einv, D.mu);
  pa::k_multi_axpy<<<grid_for(ndof), 256, 0, s>>>(y, Y, D.k, ndof, D.mu, sign, ndof);
  PA_HIP(hipGetLastError());
  return PA_OK;
}

// Jacobi(3x3 block)-PCG in FP64 on the handle's graph, on the workspace W; with a
// deflation space D, deflated PCG (the start corrected by V E^-1 V^T r, every direction
// projected: p -= V E^-1 (A V)^T z).
static int pcg_block_jacobi(pa_handle h, const double* vals, double* x, const double* b,
                            const pa_solver_opts& opts, pa_solver_stats* st, const PcgWork& W,
                            const Deflation* D = nullptr) {
  const int nn = h->nn;
  const size_t ndof = 3 * (size_t)nn;
  hipStream_t s = h->stream;
  const int G = pa::red_grid((size_t)nn, 256);       // node-parallel kernels
  const int Gs = pa::red_grid((size_t)nn, 16);       // row-group kernels (16 rows per group)
  double* P = W.partials;
  auto PS = [&](int slot) { return P + (size_t)slot * pa::kRedBlocks; };
  pa::k_diag_inv<<<grid_for(nn), 256, 0, s>>>(vals, h->diag_slot.p, W.dinv, nn);
  if (opts.use_x0) {
    pa::k_spmv<<<Gs, 256, 0, s>>>(vals, h->rowmap.p, h->colidx.p, x, b, W.r, nn, 1);
  } else {
    pa::k_fill<<<grid_for(ndof), 256, 0, s>>>(x, ndof, 0.0);
    PA_HIP(hipMemcpyAsync(W.r, b, ndof * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  if (D) {  // x += V E^-1 V^T r, r -= A V E^-1 V^T r (same mu)
    if (int rc = deflate_apply(h, *D, D->V, W.r, x, D->V, 1.0, G)) return rc;
    pa::k_multi_axpy<<<grid_for(ndof), 256, 0, s>>>(W.r, D->AV, D->k, ndof, D->mu, -1.0, ndof);
  }
  pa::k_pcg_init<<<G, 256, 0, s>>>(W.r, W.z, W.p, b, W.dinv, nn, PS(pa::kPRZ0), PS(pa::kPRR),
                                    PS(pa::kPBB));
  if (D)
    if (int rc = deflate_apply(h, *D, D->AV, W.z, W.p, D->V, -1.0, G)) return rc;
  pa::k_sum<<<1, 256, 0, s>>>(PS(pa::kPRR), G, W.scal, pa::kRR);
  pa::k_sum<<<1, 256, 0, s>>>(PS(pa::kPBB), G, W.scal, pa::kBB);
  PA_HIP(hipGetLastError());
  if (int rc = read_scal(h, W.scal)) return rc;
  const double bb = h->host_scal[pa::kBB];
  double rel = bb > 0 ? std::sqrt(h->host_scal[pa::kRR] / bb) : 0.0;
  int it = 0;
  bool conv = bb == 0.0 || rel <= opts.tolerance;
  if (bb == 0.0) pa::k_fill<<<grid_for(ndof), 256, 0, s>>>(x, ndof, 0.0);  // b = 0: the solution is 0, whatever x0
  while (!conv && it < opts.max_iters) {
    // rz of this iteration in P_RZ[cur]; the update writes the next one into P_RZ[cur ^ 1]
    const int cur = it & 1;
    pa::k_pcg_spmv<<<Gs, 256, 0, s>>>(vals, h->rowmap.p, h->colidx.p, W.p, W.q, nn, PS(pa::kPPQ));
    pa::k_pcg_update<<<G, 256, 0, s>>>(x, W.r, W.z, W.p, W.q, W.dinv, nn, PS(pa::kPPQ),
                                       Gs, PS(pa::kPRZ0 + cur), G, PS(pa::kPRZ0 + (cur ^ 1)), PS(pa::kPRR),
                                       W.scal);
    pa::k_pcg_pupdate<<<G, 256, 0, s>>>(W.p, W.z, nn, PS(pa::kPRZ0 + (cur ^ 1)), PS(pa::kPRZ0 + cur), G);
    if (D)
      if (int rc = deflate_apply(h, *D, D->AV, W.z, W.p, D->V, -1.0, G)) return rc;
    ++it;
    if (it % opts.check_every == 0 || it == opts.max_iters) {
      pa::k_sum<<<1, 256, 0, s>>>(PS(pa::kPRR), G, W.scal, pa::kRR);
      PA_HIP(hipGetLastError());
      if (int rc = read_scal(h, W.scal)) return rc;
      rel = std::sqrt(h->host_scal[pa::kRR] / bb);
      if (!std::isfinite(rel)) return fail(PA_ERR_NOT_CONVERGED, "PCG produced a non-finite residual");
      conv = rel <= opts.tolerance;
    }
  }
  // true residual
  pa::k_spmv<<<Gs, 256, 0, s>>>(vals, h->rowmap.p, h->colidx.p, x, b, W.r, nn, 1);
  pa::k_dot<<<G, 256, 0, s>>>(W.r, W.r, ndof, PS(pa::kPDot));
  pa::k_sum<<<1, 256, 0, s>>>(PS(pa::kPDot), G, W.scal, pa::kDot0);
  PA_HIP(hipGetLastError());
  if (int rc = read_scal(h, W.scal)) return rc;
  const double true_rel = bb > 0 ? std::sqrt(h->host_scal[pa::kDot0] / bb) : 0.0;
  if (st) {
    st->iterations = it;
    st->outer_iterations = 1;
    st->converged = conv ? 1 : 0;
    st->rel_residual = true_rel;
  }
  if (!conv)
    return fail(PA_ERR_NOT_CONVERGED, "PCG reached max_iters=" + std::to_string(opts.max_iters) +
                                          " at relative residual " + std::to_string(rel));
  return PA_OK;
}

int pa_solve(pa_handle h, const double* vals, double* x, const double* b, const pa_solver_opts* o,
             pa_solver_stats* st) {
  if (int rc = ready(h)) return rc;
  if (!vals || !x || !b) return fail(PA_ERR_ARG, "pa_solve: null pointer");
  pa_solver_opts opts;
  if (int rc = solve_opts(o, opts, "pa_solve")) return rc;
  if (int rc = ensure_solver_ws(h)) return rc;
  return pcg_block_jacobi(h, vals, x, b, opts, st,
                          PcgWork{h->w_r.p, h->w_z.p, h->w_p.p, h->w_q.p, h->w_dinv.p, h->partials.p, h->scal.p});
}

// ------------------------------------------------------------ solve without allocation
// Workspace of pa_solve_noalloc: a file-scope device array, one copy per GPU, sized for
// meshes of up to PA_NOALLOC_MAX_NODES nodes.  Host code never uses the array's name as an
// address (on the host it is a stand-in symbol, not device memory): the device address of
// each GPU's copy comes from hipGetSymbolAddress, taken once per device with that device
// current, and only that address is handed to kernels.
constexpr size_t kNoAllocPartials = (size_t)pa::kNumPartialSlots * pa::kRedBlocks;
constexpr size_t kNoAllocDoubles = 21 * (size_t)PA_NOALLOC_MAX_NODES + kNoAllocPartials + pa::kNumScal;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kNoAllocMaxDevices] = {};

static int noalloc_scratch(double** out) {
  int dev = 0;
  PA_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) return fail(PA_ERR_UNSUPPORTED, "pa_solve_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    PA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(PA_ERR_HIP, "pa_solve_noalloc: no device address for the workspace");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return PA_OK;
}

int pa_solve_noalloc(pa_handle h, const double* vals, double* x, const double* b, const pa_solver_opts* o,
                     pa_solver_stats* st) {
  if (int rc = ready(h)) return rc;
  if (!vals || !x || !b) return fail(PA_ERR_ARG, "pa_solve_noalloc: null pointer");
  if ((size_t)h->nn > (size_t)PA_NOALLOC_MAX_NODES)
    return fail(PA_ERR_UNSUPPORTED, "pa_solve_noalloc: " + std::to_string(h->nn) + " nodes exceed its workspace of " +
                                        std::to_string(PA_NOALLOC_MAX_NODES) + " nodes (use pa_solve)");
  pa_solver_opts opts;
  if (int rc = solve_opts(o, opts, "pa_solve_noalloc")) return rc;
  double* ws = nullptr;
  if (int rc = noalloc_scratch(&ws)) return rc;
  const size_t nd = 3 * (size_t)h->nn;
  PcgWork W{ws, ws + nd, ws + 2 * nd, ws + 3 * nd, ws + 4 * nd, ws + 4 * nd + 9 * (size_t)h->nn, nullptr};
  W.scal = W.partials + kNoAllocPartials;
  return pcg_block_jacobi(h, vals, x, b, opts, st, W);
}

// Mixed precision: FP64 iterative refinement around an FP32 matrix-free Jacobi-PCG.
// Outer: r = f - K u in FP64 (element residual kernel, constrained dofs masked);
// inner: K d = r to inner_tolerance with y = K x in FP32; u += d in FP64.
// b == nullptr: right-hand side is the traction load f; else b (device, 3N; constrained
// entries ignored).
// workspace of the matrix-free paths and the FP64 block-diagonal inverse of K(z)
static int mixed_setup(pa_handle h, const double* z) {
  const int nn = h->nn;
  const size_t ndof = 3 * (size_t)nn;
  hipStream_t s = h->stream;
  if (h->m_r.n != ndof) {
    PA_HIP(h->m_xz.alloc(nn));
    PA_HIP(h->m_dinv.alloc(9 * (size_t)nn));
    PA_HIP(h->m_r.alloc(ndof));
    PA_HIP(h->m_z.alloc(ndof));
    PA_HIP(h->m_p.alloc(ndof));
    PA_HIP(h->m_p2.alloc(ndof));
    PA_HIP(h->m_q.alloc(ndof));
    PA_HIP(h->m_R.alloc(ndof));
  }
  PA_HIP(pa::launch_make_xz(h->X.p, z, h->m_xz.p, nn, s));
  pa::DiagArgs da{};
  da.chunk_rows = h->chunk_rows.p;
  da.node_offs = h->chunk_node_offs.p;
  da.nodes = h->chunk_nodes.p;
  da.ent_ln = h->ent_ln.p;
  da.list_offs = h->list_offs.p;
  da.nchunks = h->chunks.count;
  da.max_rows = h->chunks.max_rows;
  da.max_nodes = h->chunks.max_nodes;
  da.X = h->X.p;
  da.z = z;
  da.bcmask = h->bcmask.p;
  da.dinv = h->m_dinv.p;
  da.m = h->material();
  da.gbound = h->chunk_gbound.p;
  if (h->box_fast && h->box_mf) {
    // the lattice FP32 operator (element weights of this density in FP32) and the block
    // diagonal from the FP64 weights, both summed in a fixed order
    if (h->box_W32.n != h->box_W.n) {
      PA_HIP(h->box_W32.alloc(h->box_W.n));
      // FP32 p at the lattice points, zero on the ghost layer, which k_boxmf_scatter never
      // writes, and on a tail of one padded plane and row past the end (k_boxmf_apply_kuhn
      // reads one point beyond the last plane; the tets there have zero weight)
      const size_t px = (size_t)h->box.nx + 1, pxy = px * ((size_t)h->box.ny + 1);
      PA_HIP(h->box_plat.alloc(h->box_zlat.n + pxy + px + 1));
      pa::k_fill<<<grid_for(2 * h->box_plat.n), 256, 0, s>>>(reinterpret_cast<double*>(h->box_plat.p),
                                                            2 * h->box_plat.n, 0.0);
      PA_HIP(hipGetLastError());
      PA_HIP(h->box_mf_part.alloc(pa::kRedBlocks));
    }
    const pa::BoxArgs a = box_args(h, nullptr, z, nullptr, nullptr);
    pa::BoxArgs a32 = a;  // FP32 weights for the operator, in Kuhn-path order for k_boxmf_apply_kuhn
    if (h->box_kuhn) a32.corners = h->box_corners_kuhn.p;
    box_weights(h, a32, h->box_W32.p);
    box_weights(h, a, a.W);
    if (h->box_res_kuhn)
      pa::k_box_diag_inv_kuhn<<<std::min(4096, ((nn + 255) / 256 + pa::kNumXcd - 1) / pa::kNumXcd * pa::kNumXcd), 256,
                                0, s>>>(a, h->bcmask.p, h->m_dinv.p);
    else
      pa::k_box_diag_inv<<<grid_for(((size_t)nn + 31) / 32, 1, 16384), 256, 0, s>>>(a, h->bcmask.p, h->m_dinv.p);
    PA_HIP(hipGetLastError());
  } else {
    PA_HIP(pa::launch_mf_diag_inv(da, s));
  }
  return PA_OK;
}

// y = K p through the FP32 matrix-free operator with the direction update fused (MfArgs):
// on box meshes that take the lattice assembly the lattice apply (FP32 element weights
// from mixed_setup, deterministic row sums), else the chunk apply.  *P and *np: the p.y
// partials it wrote.
static int mf_apply(pa_handle h, const pa::MfArgs& ma, const double** P, int* np) {
  hipStream_t s = h->stream;
  if (h->box_fast && h->box_mf) {
    pa::BoxMfArgs b{};
    b.gpack = h->box_gpack.p;
    b.S = h->box_S.p;
    b.W = h->box_W32.p;
    b.B = h->box_Bf.p;
    b.bcmask = h->bcmask.p;
    b.z = ma.z;
    b.p_prev = ma.p_prev;
    b.p_cur = ma.p_cur;
    b.first = ma.first;
    b.scal = ma.scal;
    b.plat = h->box_plat.p;
    b.y = ma.y;
    b.partials = h->box_mf_part.p;
    b.nrows = h->nn;
    b.nx = h->box.nx - 1;
    b.ny = h->box.ny - 1;
    b.nz = h->box.nz - 1;
    const int nblk = (h->nn + pa::kBoxRows - 1) / pa::kBoxRows;
    int G = std::min((nblk + pa::kNumXcd - 1) / pa::kNumXcd * pa::kNumXcd, h->box_grid_mf);
    pa::k_boxmf_scatter<<<grid_for((size_t)h->nn, 256, 16384), 256, 0, s>>>(b);
    if (h->box_kuhn) {
      const pa::Material m = h->material();
      for (int k = 0; k < 3; ++k) b.ih[k] = (float)(1.0 / h->box_step[k]);
      b.vol = (float)(h->box_step[0] * h->box_step[1] * h->box_step[2] / 6.0);
      b.lam = (float)m.c01;
      b.mu = (float)m.c33;
      G = std::min(pa::kRedBlocks, ((h->nn + 255) / 256 + pa::kNumXcd - 1) / pa::kNumXcd * pa::kNumXcd);
      pa::k_boxmf_apply_kuhn<<<G, 256, 0, s>>>(b);
    } else {
      pa::k_boxmf_apply<<<G, 256, 0, s>>>(b);
    }
    PA_HIP(hipGetLastError());
    *P = h->box_mf_part.p;
    *np = G;
    return PA_OK;
  }
  PA_HIP(pa::launch_mf_apply(ma, s));
  *P = h->chunk_partials.p;
  *np = h->chunks.count;
  return PA_OK;
}

static int solve_mixed(pa_handle h, const double* z, double* u, const double* b, const pa_solver_opts& opts,
                       pa_solver_stats* st) {
  const int nn = h->nn;
  const size_t ndof = 3 * (size_t)nn;
  hipStream_t s = h->stream;
  if (int rc = mixed_setup(h, z)) return rc;
  if (!opts.use_x0) pa::k_fill<<<grid_for(ndof), 256, 0, s>>>(u, ndof, 0.0);
  const int nbc = (int)h->bc_dofs.size();
  // prescribed values for the state; homogeneous constraints for a caller right-hand side
  if (nbc) pa::k_bc_set<<<grid_for(nbc), 256, 0, s>>>(u, h->bc_d.p, h->bc_v.p, nbc, b ? 0.0 : 1.0);
  PA_HIP(hipGetLastError());
  const int G = pa::red_grid((size_t)nn, 256);
  double* P = h->partials.p;
  auto PS = [&](int slot) { return P + (size_t)slot * pa::kRedBlocks; };
  pa::MfArgs ma{};
  ma.chunk_rows = h->chunk_rows.p;
  ma.node_offs = h->chunk_node_offs.p;
  ma.nodes = h->chunk_nodes.p;
  ma.ent_ln = h->ent_ln.p;
  ma.list_offs = h->list_offs.p;
  ma.nchunks = h->chunks.count;
  ma.max_rows = h->chunks.max_rows;
  ma.max_nodes = h->chunks.max_nodes;
  ma.Xz = h->m_xz.p;
  ma.bcmask = h->bcmask.p;
  ma.z = h->m_z.p;
  ma.y = h->m_q.p;
  ma.scal = h->scal.p;
  ma.partials = h->chunk_partials.p;
  ma.m = h->material();
  ma.gbound = h->chunk_gbound.p;
  // r = mask(b - R) into m_r and ||r||^2
  auto masked_residual = [&](const double* R, double* out) -> int {
    PA_HIP(pa::launch_neg_mask(R, b, h->bcmask.p, h->m_r.p, nn, PS(pa::kPDot), G, s));
    pa::k_sum<<<1, 256, 0, s>>>(PS(pa::kPDot), G, h->scal.p, pa::kDot1);
    PA_HIP(hipGetLastError());
    if (int rc = read_scal(h)) return rc;
    *out = h->host_scal[pa::kDot1];
    return PA_OK;
  };
  double bb = h->f_free_norm2;
  if (b) {  // ||b||^2 over free dofs, through the masked conversion of b - K 0
    pa::k_fill<<<grid_for(ndof), 256, 0, s>>>(h->m_R.p, ndof, 0.0);
    if (int rc = masked_residual(h->m_R.p, &bb)) return rc;
  }
  // Jacobi-PCG in FP64 whose operator K p is applied in FP32 (matrix-free); the
  // iterated residual is replaced by the true FP64 residual f - K u whenever it has
  // dropped by `delta` since the last replacement (the search direction is kept), and
  // convergence is judged on that true residual.
  int inner_total = 0, outer = 0;
  double rel = 0.0;
  bool conv = false;
  const double delta = std::min(0.5, std::max(1e-6, opts.inner_tolerance > 0 ? opts.inner_tolerance : 0.1));
  double* pbuf[2] = {h->m_p.p, h->m_p2.p};
  // true residual of u into m_r and its norm
  auto true_residual = [&](double* rr) -> int {
    if (int rc = residual(h, u, z, h->m_R.p, b == nullptr)) return rc;
    return masked_residual(h->m_R.p, rr);
  };
  double rr_ref = 0.0;
  if (int rc = true_residual(&rr_ref)) return rc;
  rel = bb > 0 ? std::sqrt(rr_ref / bb) : 0.0;
  if (!std::isfinite(rel)) return fail(PA_ERR_NOT_CONVERGED, "mixed-precision solve: non-finite residual");
  conv = bb == 0.0 || rel <= opts.tolerance;
  if (bb == 0.0) pa::k_fill<<<grid_for(ndof), 256, 0, s>>>(u, ndof, 0.0);  // f = 0: the solution is 0, whatever u0
  if (!conv) {
    PA_HIP(pa::launch_pcgm_init(h->m_r.p, h->m_z.p, h->m_dinv.p, nn, PS(pa::kPRZ0), G, s));
    PA_HIP(pa::launch_pcgm_beta(PS(pa::kPRZ0), nullptr, G, h->scal.p, pa::kBetaInit, s));
  }
  int it = 0;
  bool restart = true;  // the next apply starts the direction from p = z
  while (!conv) {
    const int cur = it & 1;
    ma.first = restart ? 1 : 0;
    restart = false;
    ma.p_prev = pbuf[cur ^ 1];
    ma.p_cur = pbuf[cur];
    const double* Pq = nullptr;
    int npq = 0;
    if (int rc = mf_apply(h, ma, &Pq, &npq)) return rc;
    PA_HIP(pa::launch_pcgm_alpha(Pq, npq, h->scal.p, s));
    PA_HIP(pa::launch_pcgm_update(u, h->m_r.p, h->m_z.p, pbuf[cur], h->m_q.p, h->m_dinv.p, nn, h->scal.p,
                                  PS(pa::kPRZ0), PS(pa::kPRR), G, s));
    PA_HIP(pa::launch_pcgm_beta(PS(pa::kPRZ0), PS(pa::kPRR), G, h->scal.p, pa::kBetaStep, s));
    ++it;
    ++inner_total;
    const bool last = inner_total >= opts.max_iters;
    if (it % opts.check_every != 0 && !last) continue;
    if (int rc = read_scal(h)) return rc;
    const double rr_it = h->host_scal[pa::kRR];
    if (!std::isfinite(rr_it)) return fail(PA_ERR_NOT_CONVERGED, "mixed-precision PCG: non-finite residual");
    // replace when the iterated residual has dropped enough, or could already meet the
    // tolerance, or at the iteration cap
    const bool could_meet = rr_it <= opts.tolerance * opts.tolerance * bb;
    if (!(rr_it <= delta * delta * rr_ref || could_meet || last)) continue;
    ++outer;
    if (int rc = true_residual(&rr_ref)) return rc;
    rel = std::sqrt(rr_ref / bb);
    if (!std::isfinite(rel)) return fail(PA_ERR_NOT_CONVERGED, "mixed-precision solve: non-finite residual");
    if (rel <= opts.tolerance) {
      conv = true;
      break;
    }
    if (last) break;
    // continue from the replaced residual: z = Dinv r, beta = (r.z) / rz of the last step;
    // restart from p = z instead when the true residual is well above the iterated one
    restart = rr_ref > 4.0 * rr_it;
    PA_HIP(pa::launch_pcgm_init(h->m_r.p, h->m_z.p, h->m_dinv.p, nn, PS(pa::kPRZ0), G, s));
    PA_HIP(pa::launch_pcgm_beta(PS(pa::kPRZ0), nullptr, G, h->scal.p, restart ? pa::kBetaInit : pa::kBetaReplace,
                                s));
  }
  if (st) {
    st->iterations = inner_total;
    st->outer_iterations = outer;
    st->converged = conv ? 1 : 0;
    st->rel_residual = rel;
  }
  if (!conv)
    return fail(PA_ERR_NOT_CONVERGED, "mixed-precision solve stopped at relative residual " + std::to_string(rel));
  return PA_OK;
}

// ---------------------------------------------------------------- geometric multigrid
// Grid coordinates (i, j, k) of every node of a box node grid with `d` nodes per axis
// over [lo, lo + len]; false when a node is off the lattice or two nodes share a point.
static bool box_grid(const double* X, int nn, const pa::BoxDims& d, const double lo[3], const double len[3],
                     std::vector<int32_t>& grid) {
  grid.assign(3 * (size_t)nn, 0);
  std::vector<uint8_t> seen((size_t)d.nx * d.ny * d.nz, 0);
  const int n3[3] = {d.nx, d.ny, d.nz};
  // Tolerance in spacings: 1e-6, or the coordinates' own round-off where that is larger
  // (lattices stored in single precision round by ~6e-8 of the largest coordinate).
  double tol[3];
  for (int a = 0; a < 3; ++a) {
    const double xmax = std::max(std::fabs(lo[a]), std::fabs(lo[a] + len[a]));
    tol[a] = std::max(1e-6, 16.0 * 6e-8 * xmax * (n3[a] - 1) / len[a]);
    if (tol[a] >= 0.25) return false;  // coordinates too coarse to resolve the lattice
  }
  for (int n = 0; n < nn; ++n) {
    for (int a = 0; a < 3; ++a) {
      const double hstep = len[a] / (n3[a] - 1);
      const double t = (X[3 * (size_t)n + a] - lo[a]) / hstep;
      const double r = std::nearbyint(t);
      if (!(std::fabs(t - r) <= tol[a]) || r < 0 || r > n3[a] - 1) return false;  // NaN fails too
      grid[3 * (size_t)n + a] = (int32_t)r;
    }
    const size_t g = ((size_t)grid[3 * (size_t)n] * d.ny + grid[3 * (size_t)n + 1]) * d.nz + grid[3 * (size_t)n + 2];
    if (seen[g]) return false;
    seen[g] = 1;
  }
  return true;
}

int pa_set_box(pa_handle h, int nx, int ny, int nz) {
  if (!h) return fail(PA_ERR_ARG, "null handle");
  if (nx < 1 || ny < 1 || nz < 1) return fail(PA_ERR_ARG, "pa_set_box: hex counts must be >= 1");
  if ((size_t)(nx + 1) * (ny + 1) * (nz + 1) != (size_t)h->nn)
    return fail(PA_ERR_MESH, "pa_set_box: node count is not (nx+1)(ny+1)(nz+1)");
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int n = 0; n < h->nn; ++n)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], h->coords[3 * (size_t)n + a]);
      hi[a] = std::max(hi[a], h->coords[3 * (size_t)n + a]);
    }
  const pa::BoxDims d{nx + 1, ny + 1, nz + 1};
  double len[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
  for (int a = 0; a < 3; ++a)
    if (!(len[a] > 0.0) || !std::isfinite(len[a]) || !std::isfinite(lo[a]))
      return fail(PA_ERR_MESH, "pa_set_box: the box has a zero-length or non-finite axis");
  std::vector<int32_t> grid;
  if (!box_grid(h->coords.data(), h->nn, d, lo, len, grid))
    return fail(PA_ERR_MESH, "pa_set_box: the nodes do not form the box's node lattice");
  h->box = d;
  for (int a = 0; a < 3; ++a) {
    h->box_lo[a] = lo[a];
    h->box_len[a] = len[a];
  }
  h->box_grid.swap(grid);
  h->has_box = true;
  h->mg.clear();
  if (h->finalized) return box_fast_setup(h);
  return PA_OK;
}

// Host replica of the element gradient formula (tet_gradients in pa_kernels.hip).
static void host_tet_gradients(const double x[4][3], double g[4][3], double& vol) {
  const double j11 = x[0][0] - x[3][0], j12 = x[1][0] - x[3][0], j13 = x[2][0] - x[3][0];
  const double j21 = x[0][1] - x[3][1], j22 = x[1][1] - x[3][1], j23 = x[2][1] - x[3][1];
  const double j31 = x[0][2] - x[3][2], j32 = x[1][2] - x[3][2], j33 = x[2][2] - x[3][2];
  const double det = j11 * j22 * j33 + j12 * j23 * j31 + j13 * j21 * j32 - j11 * j23 * j32 -
                     j12 * j21 * j33 - j13 * j22 * j31;
  const double inv = 1.0 / det;
  g[0][0] = (j22 * j33 - j23 * j32) * inv;
  g[0][1] = (j13 * j32 - j12 * j33) * inv;
  g[0][2] = (j12 * j23 - j13 * j22) * inv;
  g[1][0] = (j23 * j31 - j21 * j33) * inv;
  g[1][1] = (j11 * j33 - j13 * j31) * inv;
  g[1][2] = (j13 * j21 - j11 * j23) * inv;
  g[2][0] = (j21 * j32 - j22 * j31) * inv;
  g[2][1] = (j12 * j31 - j11 * j32) * inv;
  g[2][2] = (j11 * j22 - j12 * j21) * inv;
  for (int d = 0; d < 3; ++d) g[3][d] = -(g[0][d] + g[1][d] + g[2][d]);
  vol = std::fabs(det) * (1.0 / 6.0);
}

// Decide whether the lattice assembly applies and build its tables.  Conditions: at most
// 1024 nodes per axis; every node within max(1e-13, 8 eps |x|max / step) spacings of its
// lattice point; every
// element inside one hex, the elements of every hex being the same six corner sets; every
// graph row at most 16 blocks, all within one lattice step of the row's node.
// PA_BOX_ASSEMBLY=0 keeps the general chunk kernel.
static int box_fast_setup(pa_handle h) {
  h->box_fast = false;
  if (const char* e = std::getenv("PA_BOX_ASSEMBLY"))
    if (std::atoi(e) == 0) return PA_OK;
  const pa::BoxDims d = h->box;
  if (d.nx > pa::kBoxMaxAxis || d.ny > pa::kBoxMaxAxis || d.nz > pa::kBoxMaxAxis) return PA_OK;
  if (h->graph.max_row_len > pa::kBoxMaxRow) return PA_OK;
  const int hx = d.nx - 1, hy = d.ny - 1, hz = d.nz - 1;
  if (hx < 1 || hy < 1 || hz < 1 || (size_t)6 * hx * hy * hz != (size_t)h->ne) return PA_OK;
  const double step[3] = {h->box_len[0] / hx, h->box_len[1] / hy, h->box_len[2] / hz};
  const std::vector<int32_t>& gr = h->box_grid;
  // Lattice positions must hold to the FP64 round-off of the coordinates themselves
  // (a linspace axis of 1023 hexes deviates by ~1e-13 spacings): the general kernel's
  // own edge vectors carry that relative error, so the lattice blocks then agree with it
  // to rounding.
  double tol[3];
  for (int a = 0; a < 3; ++a) {
    const double xmax = std::max(std::fabs(h->box_lo[a]), std::fabs(h->box_lo[a] + h->box_len[a]));
    tol[a] = std::max(1e-13, 8.0 * DBL_EPSILON * xmax / step[a]);
  }
  for (int n = 0; n < h->nn; ++n)
    for (int a = 0; a < 3; ++a) {
      const double t = (h->coords[3 * (size_t)n + a] - h->box_lo[a]) / step[a];
      if (!(std::fabs(t - gr[3 * (size_t)n + a]) <= tol[a])) return PA_OK;
    }
  // element types: sorted corner ids (bit 0 x, bit 1 y, bit 2 z) relative to the hex
  std::vector<std::array<uint8_t, 4>> types;
  std::vector<uint8_t> seen((size_t)hx * hy * hz, 0);
  for (int e = 0; e < h->ne; ++e) {
    int lo[3] = {1 << 30, 1 << 30, 1 << 30};
    for (int a = 0; a < 4; ++a)
      for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], gr[3 * (size_t)h->conn[4 * (size_t)e + a] + k]);
    std::array<uint8_t, 4> cs;
    for (int a = 0; a < 4; ++a) {
      int c = 0;
      for (int k = 0; k < 3; ++k) {
        const int o = gr[3 * (size_t)h->conn[4 * (size_t)e + a] + k] - lo[k];
        if (o < 0 || o > 1) return PA_OK;
        c |= o << k;
      }
      cs[a] = (uint8_t)c;
    }
    std::sort(cs.begin(), cs.end());
    if (cs[0] == cs[1] || cs[1] == cs[2] || cs[2] == cs[3]) return PA_OK;
    if (lo[0] >= hx || lo[1] >= hy || lo[2] >= hz) return PA_OK;
    int t = 0;
    while (t < (int)types.size() && types[t] != cs) ++t;
    if (t == (int)types.size()) {
      if (types.size() == 6) return PA_OK;
      types.push_back(cs);
    }
    uint8_t& m = seen[((size_t)lo[2] * hy + lo[1]) * hx + lo[0]];
    if (m & (1u << t)) return PA_OK;
    m |= (uint8_t)(1u << t);
  }
  if (types.size() != 6) return PA_OK;
  for (uint8_t m : seen)
    if (m != 0x3f) return PA_OK;
  const pa::Graph& g = h->graph;
  // unit-weight blocks of each lattice tet: Pt[t][a][b][i][j] = vol g_a[i] g_b[j]
  std::vector<double> Pt(6 * 4 * 4 * 9);
  std::vector<uint8_t> corners(6 * 4);
  for (int t = 0; t < 6; ++t) {
    double x[4][3], gg[4][3], vol;
    for (int a = 0; a < 4; ++a) {
      corners[4 * t + a] = types[t][a];
      for (int k = 0; k < 3; ++k) x[a][k] = ((types[t][a] >> k) & 1) * step[k];
    }
    host_tet_gradients(x, gg, vol);
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b)
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) Pt[9 * ((t * 4 + a) * 4 + b) + 3 * i + j] = vol * gg[a][i] * gg[b][j];
  }
  // Stencil: the element terms of K_rc per lattice offset c - r.  The row node is corner
  // hc of the hex whose lower corner is its lattice point minus the bits of hc; in tet t
  // of that hex it is local vertex a, the column node local vertex b.
  const size_t nhex = (size_t)(hx + 2) * (hy + 2) * (hz + 2);
  if (6 * nhex >= (size_t)1 << 32) return PA_OK;  // 32-bit element-weight offsets
  const int px = hx + 2, py = hy + 2;
  std::vector<std::vector<std::array<int, 4>>> terms(27);  // (hc, t, a, b) per offset
  for (int hc = 0; hc < 8; ++hc)
    for (int t = 0; t < 6; ++t)
      for (int a = 0; a < 4; ++a) {
        if (types[t][a] != hc) continue;
        for (int b = 0; b < 4; ++b) {
          if (b == a) continue;
          const int cb = types[t][b];
          int di = 0, mul = 1;
          for (int k = 0; k < 3; ++k, mul *= 3) di += (((cb >> k) & 1) - ((hc >> k) & 1) + 1) * mul;
          terms[di].push_back({hc, t, a, b});
        }
      }
  pa::BoxStencil st{};
  std::vector<double> P(16 * pa::kBoxMaxTerms * 9, 0.0);
  int nst = 0;
  for (int di = 0; di < 27; ++di) {
    if (terms[di].empty()) continue;
    if (di == 13 || nst >= pa::kBoxDiagLane || terms[di].size() > (size_t)pa::kBoxMaxTerms) return PA_OK;
    const int l = nst++;
    st.count[l] = (int)terms[di].size();
    st.d[l][0] = (int8_t)(di % 3 - 1);
    st.d[l][1] = (int8_t)((di / 3) % 3 - 1);
    st.d[l][2] = (int8_t)(di / 9 - 1);
    st.uoff[l] = st.d[l][0] + px * (st.d[l][1] + py * st.d[l][2]);
    for (size_t e = 0; e < terms[di].size(); ++e) {
      const auto& tm = terms[di][e];
      const int hc = tm[0], t = tm[1], a = tm[2], b = tm[3];
      st.off[l][e] = t - 6 * ((hc & 1) + px * ((hc >> 1) & 1) + px * py * (hc >> 2));
      for (int q = 0; q < 9; ++q) P[9 * (l * pa::kBoxMaxTerms + e) + q] = Pt[9 * ((t * 4 + a) * 4 + b) + q];
    }
  }
  // every graph row holds exactly its node and its present stencil neighbours; the slot
  // word of a row holds the CRS position of each stencil lane's block (4 bits per lane,
  // lane 15 the diagonal)
  std::vector<int32_t> gpack(h->nn), nog((size_t)d.nx * d.ny * d.nz);
  for (int n = 0; n < h->nn; ++n) {
    const int i = gr[3 * (size_t)n], j = gr[3 * (size_t)n + 1], k = gr[3 * (size_t)n + 2];
    gpack[n] = i | j << 10 | k << 20;
    nog[i + (size_t)d.nx * (j + (size_t)d.ny * k)] = n;
  }
  std::vector<unsigned long long> perm(h->nn, 0ull);
  for (int r = 0; r < g.num_rows; ++r) {
    const int s0 = g.rowmap[r], len = g.rowmap[r + 1] - s0;
    if (len > pa::kBoxMaxRow) return PA_OK;
    int found = 0;
    auto place = [&](int node, int lane) {
      for (int s = 0; s < len; ++s)
        if (g.colidx[s0 + s] == node) {
          perm[r] |= (unsigned long long)s << (4 * lane);
          ++found;
          return true;
        }
      return false;
    };
    if (!place(r, pa::kBoxDiagLane)) return PA_OK;
    for (int l = 0; l < nst; ++l) {
      const int i = gr[3 * (size_t)r] + st.d[l][0], j = gr[3 * (size_t)r + 1] + st.d[l][1],
                k = gr[3 * (size_t)r + 2] + st.d[l][2];
      if (i < 0 || j < 0 || k < 0 || i >= d.nx || j >= d.ny || k >= d.nz) continue;
      if (!place(nog[i + (size_t)d.nx * (j + (size_t)d.ny * k)], l)) return PA_OK;
    }
    if (found != len) return PA_OK;  // (graph rows hold no repeated columns)
  }
  // load-list entry per node (the list holds the nodes with a nonzero load, ascending)
  std::vector<int32_t> lidx(h->nn, -1);
  for (int n = 0, k = 0; n < h->nn; ++n) {
    const double* fn = &h->f[3 * (size_t)n];
    if (fn[0] != 0.0 || fn[1] != 0.0 || fn[2] != 0.0) lidx[n] = k++;
  }
  PA_HIP(h->box_lidx.upload(lidx));
  PA_HIP(h->box_B.alloc(P.size()));
  PA_HIP(h->box_Bf.alloc(P.size()));
  h->box_lam = h->box_mu = 0.0;
  PA_HIP(h->box_gpack.upload(gpack));
  PA_HIP(h->box_perm.upload(perm));
  PA_HIP(h->box_corners.upload(corners));
  PA_HIP(h->box_P.upload(P));
  PA_HIP(h->box_S.upload(std::vector<pa::BoxStencil>{st}));
  PA_HIP(h->box_W.alloc((size_t)6 * (hx + 2) * (hy + 2) * (hz + 2)));
  {
    // u at the lattice points, zero on the ghost layer (k_box_scatter writes only the
    // nodes) and on a tail of one padded plane and row past the end, which
    // k_box_res_kuhn reads at weight zero
    const size_t npts = (size_t)(hx + 2) * (hy + 2) * (hz + 2), tail = (size_t)(hx + 2) * (hy + 2) + (hx + 2) + 1;
    PA_HIP(h->box_ulat.alloc(3 * (npts + tail)));
    pa::k_fill<<<grid_for(h->box_ulat.n), 256, 0, h->stream>>>(h->box_ulat.p, h->box_ulat.n, 0.0);
    PA_HIP(hipGetLastError());
    PA_HIP(hipStreamSynchronize(h->stream));  // done before any later stream (pa_set_stream) writes u
  }
  PA_HIP(h->box_zlat.alloc((size_t)(hx + 2) * (hy + 2) * (hz + 2)));
  {  // persistent grids: the resident work-groups per CU of each variant, times the CUs
    int dev = 0, cus = 256, bk = 4, br = 3;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0) {
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bk, pa::k_box_jac<false>, 256, 0) != hipSuccess || bk < 1)
        bk = 4;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&br, pa::k_box_jac<true>, 256, 0) != hipSuccess || br < 1)
        br = 3;
    }
    int bm = 4;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bm, pa::k_boxmf_apply, 256, 0) != hipSuccess || bm < 1) bm = 4;
    h->box_grid_mf = std::min(pa::kRedBlocks, (bm * cus + pa::kNumXcd - 1) / pa::kNumXcd * pa::kNumXcd);
    int bres = 4;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bres, pa::k_box_res, 256, 0) != hipSuccess || bres < 1) bres = 4;
    h->box_grid_res = (bres * cus + pa::kNumXcd - 1) / pa::kNumXcd * pa::kNumXcd;
    h->box_grid_k = (bk * cus + pa::kNumXcd - 1) / pa::kNumXcd * pa::kNumXcd;
    h->box_grid_r = (br * cus + pa::kNumXcd - 1) / pa::kNumXcd * pa::kNumXcd;
    if (const char* e = std::getenv("PA_BOX_BLOCKS"))
      if (std::atoi(e) >= pa::kNumXcd) h->box_grid_k = h->box_grid_r = std::atoi(e) / pa::kNumXcd * pa::kNumXcd;
  }
  if (const char* e = std::getenv("PA_BOX_MF")) h->box_mf = std::atoi(e) != 0;
  if (const char* e = std::getenv("PA_BOX_WS")) h->box_ws = std::atoi(e) != 0;
  if (const char* e = std::getenv("PA_BOX_RES")) h->box_res = std::atoi(e) != 0;
  h->box_seg_cap = pa::kBoxRows * g.max_row_len;
  {
    const size_t lds = pa::box_jac_ws_lds(h->box_seg_cap);
    int dev = 0, cus = 256, bw = 2;
    // the dynamic LDS of the specialised kernel exceeds the default 64 KiB limit
    for (const void* f : {reinterpret_cast<const void*>(&pa::k_box_jac_ws<true>),
                          reinterpret_cast<const void*>(&pa::k_box_jac_ws<false>),
                          reinterpret_cast<const void*>(&pa::k_box_jac_ws_ahead<true>),
                          reinterpret_cast<const void*>(&pa::k_box_jac_ws_ahead<false>)})
      if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) {
        (void)hipGetLastError();  // leave no error behind for the caller's next HIP call
        h->box_ws = false;
      }
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
      cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bw, pa::k_box_jac_ws<true>, pa::kBoxWsThreads, lds) != hipSuccess || bw < 1) {
      (void)hipGetLastError();
      bw = 1;
    }
    h->box_grid_ws = (bw * cus + pa::kNumXcd - 1) / pa::kNumXcd * pa::kNumXcd;
  }
  {
    // measured at 256x128x128 (314 MB of weights and u): 1.65 ms against 1.78 ms; at
    // 128x64x64 (39 MB, cache-resident) the plain kernel is faster (0.17 against 0.21 ms)
    const size_t lattice_bytes = (size_t)(6 + 3) * sizeof(double) * (hx + 2) * (hy + 2) * (hz + 2);
    h->box_ahead = lattice_bytes > kBoxAheadBytes;
    if (const char* e = std::getenv("PA_BOX_AHEAD")) h->box_ahead = std::atoi(e) != 0;
  }
  {
    // the six tets of every hex are the main-diagonal Kuhn paths (sorted corner sets
    // 0, e_a, e_a + e_b, 7)
    bool kuhn = true;
    std::vector<uint8_t> ck(6 * 4);
    for (int q = 0; q < 6 && kuhn; ++q) {
      const int a = pa::kKuhnPath[q][0], b = pa::kKuhnPath[q][1];
      const std::array<uint8_t, 4> want = {0, (uint8_t)(1 << a), (uint8_t)((1 << a) | (1 << b)), 7};
      const auto it = std::find(types.begin(), types.end(), want);
      if (it == types.end()) kuhn = false;
      else h->box_tperm[q] = (int)(it - types.begin());
      for (int v = 0; v < 4; ++v) ck[4 * q + v] = want[v];
    }
    h->box_kuhn = h->box_res_kuhn = kuhn;
    if (const char* e = std::getenv("PA_BOX_MF_KUHN"))
      if (std::atoi(e) == 0) h->box_kuhn = false;
    if (const char* e = std::getenv("PA_BOX_RES_KUHN"))
      if (std::atoi(e) == 0) h->box_res_kuhn = false;
    if (h->box_kuhn) PA_HIP(h->box_corners_kuhn.upload(ck));
    for (int k = 0; k < 3; ++k) h->box_step[k] = step[k];
  }
  h->box_fast = true;
  return PA_OK;
}

int pa_box_assembly(pa_handle h, int* active) {
  if (!h || !active) return fail(PA_ERR_ARG, "pa_box_assembly: null argument");
  *active = h->box_fast ? 1 : 0;
  return PA_OK;
}

// Arguments of the lattice kernels; the unit-weight material blocks (FP64 and FP32) are
// rebuilt when the material changed.
static pa::BoxArgs box_args(pa_handle h, const double* u, const double* z, double* vals, double* R) {
  pa::BoxArgs a{};
  a.rowmap = h->rowmap.p;
  a.col = h->colidx.p;
  a.gpack = h->box_gpack.p;
  a.perm = h->box_perm.p;
  a.corners = h->box_corners.p;
  a.z = z;
  a.W = h->box_W.p;
  a.B = h->box_B.p;
  a.S = h->box_S.p;
  a.u = u;
  a.ulat = R ? h->box_ulat.p : nullptr;
  for (int q = 0; q < 6; ++q) a.tperm[q] = h->box_tperm[q];
  for (int k = 0; k < 3; ++k) a.ih[k] = 1.0 / h->box_step[k];
  a.vol = h->box_step[0] * h->box_step[1] * h->box_step[2] / 6.0;
  a.zlat = h->box_zlat.p;
  a.lidx = h->box_lidx.p;
  a.load_f = h->load_f.p;
  a.vals = vals;
  a.R = R;
  a.nrows = h->nn;
  a.nx = h->box.nx - 1;
  a.ny = h->box.ny - 1;
  a.nz = h->box.nz - 1;
  a.m = h->material();
  if (a.m.c01 != h->box_lam || a.m.c33 != h->box_mu) {
    pa::k_box_material<<<1, 256, 0, h->stream>>>(h->box_P.p, h->box_B.p, h->box_Bf.p, a.m.c01, a.m.c33);
    h->box_lam = a.m.c01;
    h->box_mu = a.m.c33;
  }
  return a;
}

// z at the lattice points (and u, when a.ulat is set), then the element weights into W.
extern "C++" template <typename T>
static void box_weights(pa_handle h, const pa::BoxArgs& a, T* W) {
  // k_box_w: one wave per run of 64 padded hexes along x
  const size_t wv = (size_t)((a.nx + 2 + 63) / 64) * (a.ny + 2) * (a.nz + 2);
  pa::k_box_scatter<<<grid_for((size_t)h->nn, 256, 16384), 256, 0, h->stream>>>(a);
  if (a.m.model == pa::kPenaltySimp && a.m.pint > 0)
    pa::k_box_w<true, T><<<grid_for(64 * wv, 256, 16384), 256, 0, h->stream>>>(a, W);
  else
    pa::k_box_w<false, T><<<grid_for(64 * wv, 256, 16384), 256, 0, h->stream>>>(a, W);
}

// K (and R = K u - f) on the lattice: element weights, then the rows.
static int box_jacobian(pa_handle h, const double* u, const double* z, double* vals, double* R) {
  const pa::BoxArgs a = box_args(h, u, z, vals, R);
  box_weights(h, a, a.W);
  const int nblk = (h->nn + pa::kBoxRows - 1) / pa::kBoxRows;
  const int tiles8 = (nblk + pa::kNumXcd - 1) / pa::kNumXcd * pa::kNumXcd;
  if (h->box_ws) {
    pa::BoxArgs aw = a;
    aw.seg_cap = h->box_seg_cap;
    const size_t lds = pa::box_jac_ws_lds(h->box_seg_cap);
    const int G = std::min(tiles8, h->box_grid_ws);
    if (h->box_ahead) {
      if (R) {
        pa::k_box_jac_ws_ahead<true><<<G, pa::kBoxWsThreads, lds, h->stream>>>(aw);
        PA_HIP(hipGetLastError());
        if (int rc = sub_load(h, R, -1.0)) return rc;  // the kernel writes R = K u
      } else {
        pa::k_box_jac_ws_ahead<false><<<G, pa::kBoxWsThreads, lds, h->stream>>>(aw);
      }
    } else if (R) {
      pa::k_box_jac_ws<true><<<G, pa::kBoxWsThreads, lds, h->stream>>>(aw);
    } else {
      pa::k_box_jac_ws<false><<<G, pa::kBoxWsThreads, lds, h->stream>>>(aw);
    }
  } else {
    const int G = std::min(tiles8, R ? h->box_grid_r : h->box_grid_k);
    if (R)
      pa::k_box_jac<true><<<G, 256, 0, h->stream>>>(a);
    else
      pa::k_box_jac<false><<<G, 256, 0, h->stream>>>(a);
  }
  PA_HIP(hipGetLastError());
  return PA_OK;
}

// The Jacobian (and optionally R = K u - f) through the lattice kernel when it applies,
// else the chunk kernel.
static int assemble_jacobian(pa_handle h, const double* u, const double* z, double* vals, double* R) {
  if (h->box_fast) return box_jacobian(h, u, z, vals, R);
  if (R) return run_chunk(h, pa::kModeJacobianResidual, u, z, vals, R, true);
  return run_chunk(h, pa::kModeJacobian, u, z, vals, nullptr);
}

static int mg_level_init(MgLevel& L, const std::vector<int32_t>& grid) {
  PA_HIP(L.grid.upload(grid));
  std::vector<int32_t> nog((size_t)L.dims.nx * L.dims.ny * L.dims.nz, 0);
  for (int n = 0; n < L.nn; ++n)
    nog[((size_t)grid[3 * (size_t)n] * L.dims.ny + grid[3 * (size_t)n + 1]) * L.dims.nz + grid[3 * (size_t)n + 2]] = n;
  PA_HIP(L.node_of_grid.upload(nog));
  const size_t nd = 3 * (size_t)L.nn;
  PA_HIP(L.r.alloc(nd));
  PA_HIP(L.x.alloc(nd));
  PA_HIP(L.y.alloc(nd));
  PA_HIP(L.d.alloc(nd));
  PA_HIP(L.t.alloc(nd));
  PA_HIP(L.rr.alloc(nd));
  PA_HIP(L.pw.alloc(nd));
  PA_HIP(L.w.alloc(L.nn));
  PA_HIP(L.dinv.alloc(9 * (size_t)L.nn));
  return PA_OK;
}

// Levels: the problem itself, then boxes with half the hexes per axis while every count
// stays even and the level has more than 64 nodes; the last level is solved densely.
static int mg_build(pa_handle h) {
  if (!h->mg.empty()) return PA_OK;
  if (!h->has_box) return fail(PA_ERR_STATE, "multigrid needs the box grid (pa_set_box)");
  // tuning knobs: PA_MG_FP32_SMOOTH=0|1, PA_MG_DEG0=<pre><post> (each 1 or 2) on level 0,
  // PA_MG_DEGC=1|2 on the coarser levels, PA_MG_RATIO = lambda_max / lambda_min,
  // PA_MG_F32=0|1 (FP32 coarse matrices in the mixed solve's smoothing)
  if (const char* e = std::getenv("PA_MG_FP32_SMOOTH")) h->mg_fp32_smooth = std::atoi(e) != 0;
  if (const char* e = std::getenv("PA_MG_DEG0")) {
    if (std::strlen(e) == 2 && (e[0] == '1' || e[0] == '2') && (e[1] == '1' || e[1] == '2')) {
      h->mg_pre0 = e[0] - '0';
      h->mg_post0 = e[1] - '0';
    }
  }
  if (const char* e = std::getenv("PA_MG_DEGC")) h->mg_degc = std::atoi(e) == 1 ? 1 : 2;
  if (const char* e = std::getenv("PA_MG_F32")) h->mg_f32 = std::atoi(e) != 0;
  if (const char* e = std::getenv("PA_MG_RATIO")) h->mg_ratio = std::max(1.5, std::atof(e));
  auto L0 = std::make_unique<MgLevel>();
  L0->h = h;
  L0->dims = h->box;
  L0->nn = h->nn;
  if (int rc = mg_level_init(*L0, h->box_grid)) return rc;
  std::vector<int32_t> prev_grid = h->box_grid;
  h->mg.push_back(std::move(L0));
  while (true) {
    MgLevel& P = *h->mg.back();
    const int hx = P.dims.nx - 1, hy = P.dims.ny - 1, hz = P.dims.nz - 1;
    if (P.nn <= 64 || hx % 2 || hy % 2 || hz % 2) break;
    pa::BoxMesh bm = pa::build_box(hx / 2, hy / 2, hz / 2, h->box_len[0], h->box_len[1], h->box_len[2]);
    for (int n = 0; n < bm.num_nodes; ++n)
      for (int a = 0; a < 3; ++a) bm.coords[3 * (size_t)n + a] += h->box_lo[a];
    auto L = std::make_unique<MgLevel>();
    L->dims = pa::BoxDims{hx / 2 + 1, hy / 2 + 1, hz / 2 + 1};
    L->nn = bm.num_nodes;
    std::vector<int32_t> grid;
    if ((size_t)bm.num_nodes != (size_t)L->dims.nx * L->dims.ny * L->dims.nz ||
        !box_grid(bm.coords.data(), bm.num_nodes, L->dims, h->box_lo, h->box_len, grid))
      return fail(PA_ERR_MESH, "multigrid: coarse box does not form its lattice");
    pa_handle hc = nullptr;
    if (int rc = pa_create(&hc, bm.num_nodes, bm.num_elems, bm.coords.data(), bm.conn.data())) return rc;
    L->h = hc;
    L->owned = true;
    if (int rc = pa_set_material(hc, h->E, h->nu)) return rc;
    // coarse levels take the homogenised stiffness weight itself as their density
    // (SIMP exponent 1, minimum 0: w = rho)
    if (int rc = pa_set_penalty(hc, 1.0, 0.0)) return rc;
    // a coarse dof is constrained when the coincident node of the finer level is
    std::vector<uint8_t> pmask(P.nn, 0);
    for (int32_t dof : P.h->bc_dofs) pmask[dof / 3] |= (uint8_t)(1u << (dof % 3));
    std::vector<int32_t> pnog((size_t)P.dims.nx * P.dims.ny * P.dims.nz, 0);
    for (int n = 0; n < P.nn; ++n)
      pnog[((size_t)prev_grid[3 * (size_t)n] * P.dims.ny + prev_grid[3 * (size_t)n + 1]) * P.dims.nz +
           prev_grid[3 * (size_t)n + 2]] = n;
    for (int comp = 0; comp < 3; ++comp) {
      std::vector<int32_t> nodes;
      for (int n = 0; n < L->nn; ++n) {
        const size_t pg = ((size_t)(2 * grid[3 * (size_t)n]) * P.dims.ny + 2 * grid[3 * (size_t)n + 1]) * P.dims.nz +
                          2 * grid[3 * (size_t)n + 2];
        if ((pmask[pnog[pg]] >> comp) & 1u) nodes.push_back(n);
      }
      if (!nodes.empty())
        if (int rc = pa_add_essential_bc(hc, (int)nodes.size(), nodes.data(), comp, 0.0)) return rc;
    }
    if (int rc = pa_finalize(hc)) return rc;
    hc->stream = h->stream;
    // the coarse box takes the lattice assembly (same values to rounding, summed in a
    // fixed order); if it does not qualify, the chunk kernel assembles it
    if (pa_set_box(hc, hx / 2, hy / 2, hz / 2) != PA_OK) hc->box_fast = false;
    if (int rc = mg_level_init(*L, grid)) return rc;
    PA_HIP(L->vals.alloc(9 * (size_t)hc->graph.rowmap[hc->nn]));
    prev_grid.swap(grid);
    h->mg.push_back(std::move(L));
  }
  if (h->mg.size() < 2) return fail(PA_ERR_UNSUPPORTED, "multigrid: the box has no coarser level");
  const size_t nc = 3 * (size_t)h->mg.back()->nn;
  PA_HIP(h->mg_Ainv.alloc(nc * nc));
  return PA_OK;
}

// y = K_l x: the fine operator (FP64 CRS `vals0`, or the FP32 matrix-free apply) on
// level 0, the assembled coarse CRS below.
static int mg_apply(pa_handle h, size_t l, const double* x, double* y, const double* vals0, bool mixed) {
  hipStream_t s = h->stream;
  if (l == 0) {
    if (mixed) {
      pa::MfArgs ma{};
      ma.chunk_rows = h->chunk_rows.p;
      ma.node_offs = h->chunk_node_offs.p;
      ma.nodes = h->chunk_nodes.p;
      ma.ent_ln = h->ent_ln.p;
      ma.list_offs = h->list_offs.p;
      ma.nchunks = h->chunks.count;
      ma.max_rows = h->chunks.max_rows;
      ma.max_nodes = h->chunks.max_nodes;
      ma.Xz = h->m_xz.p;
      ma.bcmask = h->bcmask.p;
      ma.z = x;
      ma.p_prev = nullptr;
      // the lattice operator needs no FP64 copy of x (constrained rows give y = 0, the
      // masked value, and the p.y partials go unused); the chunk apply writes one
      ma.p_cur = h->box_fast && h->box_mf ? nullptr : h->g_pc.p;
      ma.first = 1;
      ma.scal = h->scal.p;
      ma.y = y;
      ma.partials = h->chunk_partials.p;
      ma.m = h->material();
      ma.gbound = h->chunk_gbound.p;
      const double* Pq = nullptr;
      int npq = 0;
      if (int rc = mf_apply(h, ma, &Pq, &npq)) return rc;
    } else {
      pa::k_spmv<<<pa::red_grid((size_t)h->nn, 16), 256, 0, s>>>(vals0, h->rowmap.p, h->colidx.p, x, nullptr, y,
                                                                  h->nn, 0);
    }
  } else {
    MgLevel& L = *h->mg[l];
    pa::k_spmv<<<pa::red_grid((size_t)L.nn, 16), 256, 0, s>>>(L.vals.p, L.h->rowmap.p, L.h->colidx.p, x, nullptr,
                                                                y, L.nn, 0);
  }
  PA_HIP(hipGetLastError());
  return PA_OK;
}

static const double* mg_dinv(pa_handle h, size_t l, bool mixed) {
  if (l > 0) return h->mg[l]->dinv.p;
  return mixed ? h->m_dinv.p : h->w_dinv.p;
}

// Fused block-CRS smoothing step on level l (l > 0, or level 0 of the FP64 solve); the
// mixed solve's coarse levels read their FP32 copy of the matrix.
static int mg_rows(pa_handle h, size_t l, int mode, const double* vals0, const double* x, const double* r,
                   double* d, double* out, double* y, const double* base, bool mixed) {
  MgLevel& L = *h->mg[l];
  pa::MgRowArgs a{};
  a.vals = l == 0 ? vals0 : L.vals.p;
  a.vals32 = l > 0 && mixed && h->mg_f32 ? L.vals32.p : nullptr;
  a.rowmap = L.h->rowmap.p;
  a.col = L.h->colidx.p;
  a.x = x;
  a.r = r;
  a.dinv = mg_dinv(h, l, false);
  a.d = d;
  a.out = out;
  a.y = y;
  a.base = base;
  a.c1 = L.c1;
  a.c2 = L.c2;
  a.inv_theta = L.inv_theta;
  a.nrows = L.nn;
  PA_HIP(pa::launch_mg_rows(mode, a, h->stream));
  return PA_OK;
}

// x = V(r): symmetric V-cycle, degree-2 Chebyshev block-Jacobi pre- and post-smoothing,
// full-weighting restriction, trilinear prolongation, dense solve on the last level.
// `started`: the caller has already set x = d = D^-1 r / theta (fused into its
// restriction).  Where the level has an assembled operator the smoothing steps are
// fused with the operator product (6 launches per level); level 0 of the mixed solve
// applies the FP32 matrix-free operator and runs the steps as separate kernels.
static int mg_vcycle(pa_handle h, size_t l, const double* r, double* x, const double* vals0, bool mixed,
                     bool started = false) {
  hipStream_t s = h->stream;
  MgLevel& L = *h->mg[l];
  const size_t nd = 3 * (size_t)L.nn;
  if (l + 1 == h->mg.size()) {
    PA_HIP(pa::launch_dense_gemv(h->mg_Ainv.p, r, x, (int)nd, s));
    return PA_OK;
  }
  MgLevel& C = *h->mg[l + 1];
  const bool c_last = l + 2 == h->mg.size();
  const double* dinv = mg_dinv(h, l, mixed);
  const bool fused = !(l == 0 && mixed);
  const int pre = l == 0 ? h->mg_pre0 : h->mg_degc, post = l == 0 ? h->mg_post0 : h->mg_degc;
  // d is read later only by a second pre-smoothing step
  if (!started) PA_HIP(pa::launch_cheb_first(x, pre == 2 ? L.d.p : nullptr, r, dinv, L.inv_theta, L.nn, s));
  double* x1 = x;  // iterate after pre-smoothing (and, later, the coarse correction)
  if (fused) {
    if (pre == 2) {
      x1 = L.pw.p;
      if (int rc = mg_rows(h, l, pa::kMgCheb, vals0, x, r, L.d.p, x1, nullptr, nullptr, mixed)) return rc;
    }
    if (int rc = mg_rows(h, l, pa::kMgRes, vals0, x1, r, nullptr, L.rr.p, nullptr, nullptr, mixed)) return rc;
  } else {
    if (pre == 2) {
      if (int rc = mg_apply(h, l, x, L.t.p, vals0, mixed)) return rc;
      PA_HIP(pa::launch_cheb_next(x, L.d.p, r, L.t.p, dinv, L.c1, L.c2, L.nn, s));
    }
    if (int rc = mg_apply(h, l, x, L.t.p, vals0, mixed)) return rc;
    // r - K x materialised: the restriction reading r and K x instead measured 2.7x slower
    PA_HIP(pa::launch_sub(L.rr.p, r, L.t.p, nd, s));
  }
  // coarse-grid correction; the coarse pre-smoothing start is fused into the restriction
  PA_HIP(pa::launch_mg_restrict(L.rr.p, C.r.p, C.grid.p, L.node_of_grid.p, L.dims, C.h->bcmask.p, C.nn, s,
                                c_last ? nullptr : C.dinv.p, C.inv_theta, c_last ? nullptr : C.x.p,
                                c_last ? nullptr : C.d.p));
  if (int rc = mg_vcycle(h, l + 1, C.r.p, C.x.p, vals0, mixed, !c_last)) return rc;
  PA_HIP(pa::launch_mg_prolong_add(C.x.p, x1, L.grid.p, C.node_of_grid.p, C.dims, L.h->bcmask.p, L.nn, s));
  // post-smoothing on the remaining residual
  if (fused) {
    if (int rc = mg_rows(h, l, pa::kMgResCheb0, vals0, x1, r, L.d.p, L.rr.p, L.y.p, nullptr, mixed)) return rc;
    if (post == 2) {
      if (int rc = mg_rows(h, l, pa::kMgChebAdd, vals0, L.y.p, L.rr.p, L.d.p, x, nullptr, x1, mixed)) return rc;
    } else {
      PA_HIP(pa::launch_add2(x, x1, L.y.p, nd, s));
    }
  } else {
    if (int rc = mg_apply(h, l, x, L.t.p, vals0, mixed)) return rc;
    if (post == 2) {
      PA_HIP(pa::launch_sub(L.rr.p, r, L.t.p, nd, s));
      PA_HIP(pa::launch_cheb_first(L.y.p, L.d.p, L.rr.p, dinv, L.inv_theta, L.nn, s));
    } else {  // one step: y = D^-1 (r - K x) / theta, d unused
      PA_HIP(pa::launch_cheb_first(L.y.p, nullptr, r, dinv, L.inv_theta, L.nn, s, L.t.p));
    }
    if (post == 2) {
      if (int rc = mg_apply(h, l, L.y.p, L.t.p, vals0, mixed)) return rc;
      PA_HIP(pa::launch_cheb_next(L.y.p, L.d.p, L.rr.p, L.t.p, dinv, L.c1, L.c2, L.nn, s));
    }
    PA_HIP(pa::launch_add(x, L.y.p, nd, s));
  }
  return PA_OK;
}

// sum of per-group partials of a.b (n3 values) into scal[slot], read back
static int mg_dot(pa_handle h, const double* a, const double* b, size_t n, int slot, double* out) {
  double* P = h->partials.p + (size_t)pa::kPDot * pa::kRedBlocks;
  const int G = pa::red_grid(n / 3 + 1, 256);
  pa::k_dot<<<G, 256, 0, h->stream>>>(a, b, n, P);
  pa::k_sum<<<1, 256, 0, h->stream>>>(P, G, h->scal.p, slot);
  PA_HIP(hipGetLastError());
  if (out) {
    if (int rc = read_scal(h)) return rc;
    *out = h->host_scal[slot];
  }
  return PA_OK;
}

// PA_MG_TRACE=1: per-level spectral estimates and the convergence checks on stderr
static bool mg_trace() {
  static const bool on = std::getenv("PA_MG_TRACE") != nullptr;
  return on;
}

// Inverse of a symmetric positive definite n x n matrix (row major, in place in A;
// U is n x n scratch): Cholesky A = L L^T into the lower triangle, U = L^-T (upper),
// A^-1 = U U^T.  Every inner loop runs along contiguous rows.
static int spd_inverse(std::vector<double>& A, std::vector<double>& U, int n) {
  for (int j = 0; j < n; ++j) {
    double* Aj = &A[(size_t)j * n];
    double d = Aj[j];
    for (int k = 0; k < j; ++k) d -= Aj[k] * Aj[k];
    if (!(d > 0.0)) return fail(PA_ERR_NOT_CONVERGED, "multigrid: coarsest operator is not positive definite");
    const double ljj = std::sqrt(d), inv = 1.0 / ljj;
    Aj[j] = ljj;
    for (int i = j + 1; i < n; ++i) {
      double* Ai = &A[(size_t)i * n];
      double t = Ai[j];
      for (int k = 0; k < j; ++k) t -= Ai[k] * Aj[k];
      Ai[j] = t * inv;
    }
  }
  // row j of U holds column j of L^-1: U[j][i] = (L^-1)[i][j], i >= j
  for (int j = 0; j < n; ++j) {
    double* Uj = &U[(size_t)j * n];
    Uj[j] = 1.0 / A[(size_t)j * n + j];
    for (int i = j + 1; i < n; ++i) {
      const double* Li = &A[(size_t)i * n];
      double t = 0.0;
      for (int k = j; k < i; ++k) t += Li[k] * Uj[k];
      Uj[i] = -t / Li[i];
    }
  }
  // A^-1[i][j] = sum_{k >= max(i, j)} (L^-1)[k][i] (L^-1)[k][j] = sum_k U[i][k] U[j][k]
  for (int i = 0; i < n; ++i) {
    const double* Ui = &U[(size_t)i * n];
    for (int j = i; j < n; ++j) {
      const double* Uj = &U[(size_t)j * n];
      double t = 0.0;
      for (int k = j; k < n; ++k) t += Ui[k] * Uj[k];
      A[(size_t)i * n + j] = t;
      A[(size_t)j * n + i] = t;
    }
  }
  return PA_OK;
}

// Chebyshev coefficients of level L for the interval [lmax / ratio, lmax].
static void mg_set_interval(pa_handle h, MgLevel& L, double lmax) {
  const double lmin = lmax / h->mg_ratio;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
  const double sigma = theta / delta, rho0 = 1.0 / sigma, rho1 = 1.0 / (2.0 * sigma - rho0);
  L.lmax = lmax;
  L.inv_theta = 1.0 / theta;
  L.c1 = rho1 * rho0;
  L.c2 = 2.0 * rho1 / delta;
}

// Largest eigenvalue of the n x n symmetric tridiagonal (diagonal a, off-diagonal b) by
// Sturm-sequence bisection.
static double tridiag_max_eig(const std::vector<double>& a, const std::vector<double>& b) {
  const size_t n = a.size();
  double lo = a[0], hi = a[0];
  for (size_t i = 0; i < n; ++i) {  // Gershgorin interval
    const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(b[i]) : 0.0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
  }
  auto count_below = [&](double x) {  // eigenvalues < x
    int c = 0;
    double q = 1.0;
    for (size_t i = 0; i < n; ++i) {
      const double bb = i > 0 ? b[i - 1] * b[i - 1] : 0.0;
      q = a[i] - x - (i > 0 ? bb / q : 0.0);
      if (q == 0.0) q = -1e-300;
      if (q < 0.0) ++c;
    }
    return c;
  };
  for (int it = 0; it < 200 && hi - lo > 1e-12 * std::max(1.0, std::fabs(hi)); ++it) {
    const double mid = 0.5 * (lo + hi);
    if (count_below(mid) == (int)n) hi = mid;
    else lo = mid;
  }
  return hi;
}

// lambda_max(D^-1 K) of level l from below: the largest Ritz value of kLanczosSteps steps
// of block-Jacobi PCG on K x = v (v a fixed pseudo-random vector), whose coefficients give
// the Lanczos tridiagonal T_jj = 1/alpha_j + beta_{j-1}/alpha_{j-1},
// T_j,j+1 = sqrt(beta_j)/alpha_j (Saad, Iterative Methods for Sparse Linear Systems, 6.7.3).
// Power iteration from the same start converges too slowly on high-contrast SIMP designs,
// whose top eigenvectors sit on a few solid/void interfaces (a 12-step power estimate of
// 2.13 against a true 2.63 made the smoother divergent there).
static int mg_lambda_max(pa_handle h, size_t l, const double* vals0, bool mixed, double* lam) {
  constexpr int kLanczosSteps = 16;
  hipStream_t s = h->stream;
  MgLevel& L = *h->mg[l];
  const size_t nd = 3 * (size_t)L.nn;
  const double* dinv = mg_dinv(h, l, mixed);
  double *r = L.r.p, *z = L.y.p, *p = L.pw.p, *q = L.t.p;
  std::vector<double> v0(nd);
  for (size_t i = 0; i < nd; ++i) v0[i] = (double)((i * 2654435761ull >> 7) % 2001) / 1000.0 - 1.0;
  PA_HIP(hipMemcpyAsync(r, v0.data(), nd * sizeof(double), hipMemcpyHostToDevice, s));
  PA_HIP(pa::launch_dinv_apply(z, dinv, r, L.nn, s));
  PA_HIP(hipMemcpyAsync(p, z, nd * sizeof(double), hipMemcpyDeviceToDevice, s));
  double rz = 0.0;
  if (int rc = mg_dot(h, r, z, nd, pa::kDot0, &rz)) return rc;
  std::vector<double> alpha, beta, ta, tb;
  const double rz0 = rz;
  for (int j = 0; j < kLanczosSteps && rz > 1e-28 * rz0; ++j) {
    if (int rc = mg_apply(h, l, p, q, vals0, mixed)) return rc;
    double pq = 0.0;
    if (int rc = mg_dot(h, p, q, nd, pa::kDot0, &pq)) return rc;
    if (!(pq > 0.0) || !std::isfinite(pq)) break;
    const double a = rz / pq;
    const size_t g = grid_for(nd);
    pa::k_axpy<<<g, 256, 0, s>>>(r, q, nd, -a);
    PA_HIP(pa::launch_dinv_apply(z, dinv, r, L.nn, s));
    double rz1 = 0.0;
    if (int rc = mg_dot(h, r, z, nd, pa::kDot0, &rz1)) return rc;
    const double b = rz1 / rz;
    ta.push_back(1.0 / a + (j > 0 ? beta.back() / alpha.back() : 0.0));
    if (j > 0) tb.push_back(std::sqrt(beta.back()) / alpha.back());
    alpha.push_back(a);
    beta.push_back(b);
    pa::k_scale<<<g, 256, 0, s>>>(p, nd, b);
    pa::k_axpy<<<g, 256, 0, s>>>(p, z, nd, 1.0);
    PA_HIP(hipGetLastError());
    rz = rz1;
  }
  if (ta.empty()) return fail(PA_ERR_NOT_CONVERGED, "multigrid: bad spectral estimate");
  *lam = tridiag_max_eig(ta, tb);
  if (!(*lam > 0.0) || !std::isfinite(*lam)) return fail(PA_ERR_NOT_CONVERGED, "multigrid: bad spectral estimate");
  return PA_OK;
}

// Per solve: homogenised coarse densities (restricted SIMP weights), coarse Jacobians,
// their block diagonals, Chebyshev intervals from Lanczos estimates, coarsest inverse.
static int mg_setup(pa_handle h, const double* z, const double* vals0, bool mixed) {
  hipStream_t s = h->stream;
  const size_t nl = h->mg.size();
  PA_HIP(pa::launch_simp_w(z, h->mg[0]->w.p, h->nn, h->material(), s));
  for (size_t l = 1; l < nl; ++l) {
    MgLevel& F = *h->mg[l - 1];
    MgLevel& L = *h->mg[l];
    PA_HIP(pa::launch_mg_wrestrict(F.w.p, L.w.p, L.grid.p, F.node_of_grid.p, F.dims, L.nn, s));
    if (int rc = assemble_jacobian(L.h, nullptr, L.w.p, L.vals.p, nullptr)) return rc;
    if (int rc = pa_apply_constraints(L.h, L.vals.p, L.rr.p, 0.0)) return rc;
    pa::k_diag_inv<<<grid_for(L.nn), 256, 0, s>>>(L.vals.p, L.h->diag_slot.p, L.dinv.p, L.nn);
    PA_HIP(hipGetLastError());
    if (mixed && h->mg_f32 && l + 1 < nl) {  // smoothed levels: the FP32 copy they stream
      if (L.vals32.n != L.vals.n) PA_HIP(L.vals32.alloc(L.vals.n));
      PA_HIP(pa::launch_to_f32(L.vals.p, L.vals32.p, L.vals.n, s));
    }
  }
  // Chebyshev interval [lmax / ratio, lmax], lmax = 1.1 x the largest Ritz value of D^-1 K
  // per smoothed level
  for (size_t l = 0; l + 1 < nl; ++l) {
    MgLevel& L = *h->mg[l];
    double lam = 0.0;
    if (int rc = mg_lambda_max(h, l, vals0, mixed, &lam)) return rc;
    L.lam = lam;
    if (mg_trace()) std::fprintf(stderr, "mg level %zu: %d nodes, lambda_max(D^-1 K) ~ %.6g\n", l, L.nn, lam);
    mg_set_interval(h, L, 1.1 * lam);
  }
  // coarsest level: dense inverse on the host (a few hundred dofs)
  MgLevel& Lc = *h->mg.back();
  const int nc = 3 * Lc.nn;
  const pa::Graph& g = Lc.h->graph;
  std::vector<double> v(9 * (size_t)g.rowmap[Lc.nn]);
  PA_HIP(hipMemcpyAsync(v.data(), Lc.vals.p, v.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  PA_HIP(hipStreamSynchronize(s));
  std::vector<double> A((size_t)nc * nc, 0.0), U((size_t)nc * nc, 0.0);
  for (int r = 0; r < Lc.nn; ++r)
    for (int k = g.rowmap[r]; k < g.rowmap[r + 1]; ++k)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[(size_t)(3 * r + i) * nc + 3 * g.colidx[k] + j] = v[9 * (size_t)k + 3 * i + j];
  if (int rc = spd_inverse(A, U, nc)) return rc;
  PA_HIP(hipMemcpyAsync(h->mg_Ainv.p, A.data(), A.size() * sizeof(double), hipMemcpyHostToDevice, s));
  PA_HIP(hipStreamSynchronize(s));
  return PA_OK;
}

// A negative r.V(r) during the multigrid PCG widens every Chebyshev interval by kMgWiden
// and restarts, at most kMgMaxWidenings times per solve.
constexpr double kMgWiden = 1.5;
constexpr int kMgMaxWidenings = 3;

// Multigrid-preconditioned CG.  FP64 path (mixed = false): the assembled, constrained
// K in `vals0`, right-hand side b, x from zero (or x0).  Mixed path: the FP32
// matrix-free operator on level 0, x = u with its prescribed values, right-hand side the
// traction load (b == nullptr) or b (homogeneous constraints), and the iterated residual
// replaced by the true FP64 residual every `inner_tolerance` decrease.
static int solve_mg(pa_handle h, const double* z, double* x, const double* b, const double* vals0,
                    pa_solver_opts opts, pa_solver_stats* st, bool mixed) {
  if (int rc = mg_build(h)) return rc;
  const int nn = h->nn;
  const size_t ndof = 3 * (size_t)nn;
  hipStream_t s = h->stream;
  for (auto& L : h->mg)
    if (L->owned) L->h->stream = s;
  if (int rc = ensure_solver_ws(h)) return rc;
  if (h->g_r.n != ndof) {
    PA_HIP(h->g_r.alloc(ndof));
    PA_HIP(h->g_z.alloc(ndof));
    PA_HIP(h->g_p.alloc(ndof));
    PA_HIP(h->g_q.alloc(ndof));
    PA_HIP(h->g_pc.alloc(ndof));
  }
  const int G = pa::red_grid((size_t)nn, 256);
  double* P = h->partials.p;
  auto PS = [&](int slot) { return P + (size_t)slot * pa::kRedBlocks; };
  double bb = 0.0, rr0 = 0.0;
  const double delta = std::min(0.5, std::max(1e-6, opts.inner_tolerance > 0 ? opts.inner_tolerance : 0.1));
  // true residual r = mask(b_eff - K x) into g_r (mixed path)
  auto true_residual = [&](double* rr) -> int {
    if (int rc = residual(h, x, z, h->m_R.p, b == nullptr)) return rc;
    PA_HIP(pa::launch_neg_mask(h->m_R.p, b, h->bcmask.p, h->g_r.p, nn, PS(pa::kPDot), G, s));
    pa::k_sum<<<1, 256, 0, s>>>(PS(pa::kPDot), G, h->scal.p, pa::kDot1);
    PA_HIP(hipGetLastError());
    if (int rc = read_scal(h)) return rc;
    *rr = h->host_scal[pa::kDot1];
    return PA_OK;
  };
  if (mixed) {
    if (int rc = mixed_setup(h, z)) return rc;
    if (!opts.use_x0) pa::k_fill<<<grid_for(ndof), 256, 0, s>>>(x, ndof, 0.0);
    const int nbc = (int)h->bc_dofs.size();
    if (nbc) pa::k_bc_set<<<grid_for(nbc), 256, 0, s>>>(x, h->bc_d.p, h->bc_v.p, nbc, b ? 0.0 : 1.0);
    PA_HIP(hipGetLastError());
    bb = h->f_free_norm2;
    if (b) {  // ||mask(b)||^2
      pa::k_fill<<<grid_for(ndof), 256, 0, s>>>(h->m_R.p, ndof, 0.0);
      PA_HIP(pa::launch_neg_mask(h->m_R.p, b, h->bcmask.p, h->g_r.p, nn, PS(pa::kPDot), G, s));
      pa::k_sum<<<1, 256, 0, s>>>(PS(pa::kPDot), G, h->scal.p, pa::kDot1);
      PA_HIP(hipGetLastError());
      if (int rc = read_scal(h)) return rc;
      bb = h->host_scal[pa::kDot1];
    }
    if (int rc = true_residual(&rr0)) return rc;
  } else {
    pa::k_diag_inv<<<grid_for(nn), 256, 0, s>>>(vals0, h->diag_slot.p, h->w_dinv.p, nn);
    if (opts.use_x0) {
      pa::k_spmv<<<pa::red_grid((size_t)nn, 16), 256, 0, s>>>(vals0, h->rowmap.p, h->colidx.p, x, b, h->g_r.p, nn, 1);
    } else {
      pa::k_fill<<<grid_for(ndof), 256, 0, s>>>(x, ndof, 0.0);
      PA_HIP(hipMemcpyAsync(h->g_r.p, b, ndof * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    PA_HIP(hipGetLastError());
    if (int rc = mg_dot(h, b, b, ndof, pa::kDot0, &bb)) return rc;
    if (int rc = mg_dot(h, h->g_r.p, h->g_r.p, ndof, pa::kDot0, &rr0)) return rc;
  }
  // level-0 smoother operator: FP32 matrix-free under the mixed solve (and, if chosen,
  // under the FP64 solve, whose CG operator stays the FP64 CRS matrix)
  const bool mf0 = mixed || h->mg_fp32_smooth;
  if (mf0 && !mixed)
    if (int rc = mixed_setup(h, z)) return rc;
  if (int rc = mg_setup(h, z, vals0, mf0)) return rc;
  double rel = bb > 0 ? std::sqrt(rr0 / bb) : 0.0;
  if (!std::isfinite(rel)) return fail(PA_ERR_NOT_CONVERGED, "multigrid PCG: non-finite residual");
  bool conv = bb == 0.0 || rel <= opts.tolerance;
  if (bb == 0.0) pa::k_fill<<<grid_for(ndof), 256, 0, s>>>(x, ndof, 0.0);  // b = 0: the solution is 0, whatever x0
  int it = 0, replacements = 0, widenings = 0;
  double rr_ref = rr0;
  // z = V(r) and rz; kBetaInit also starts p = z, kBetaReplace recomputes beta for a
  // replaced residual (the direction update p = z + beta p follows in the loop)
  auto precondition = [&](int mode) -> int {
    if (int rc = mg_vcycle(h, 0, h->g_r.p, h->g_z.p, vals0, mf0)) return rc;
    pa::k_dot<<<G, 256, 0, s>>>(h->g_r.p, h->g_z.p, ndof, PS(pa::kPRZ0));
    PA_HIP(pa::launch_pcgm_beta(PS(pa::kPRZ0), mode == pa::kBetaStep ? PS(pa::kPRR) : nullptr, G, h->scal.p, mode,
                                s));
    if (mode == pa::kBetaInit)
      PA_HIP(hipMemcpyAsync(h->g_p.p, h->g_z.p, ndof * sizeof(double), hipMemcpyDeviceToDevice, s));
    return PA_OK;
  };
  if (!conv)
    if (int rc = precondition(pa::kBetaInit)) return rc;
  // a multigrid iteration reduces the residual several times over: check (and, mixed,
  // replace the residual) at least every second iteration
  const int check_every = std::min(opts.check_every, 2);
  while (!conv && it < opts.max_iters) {
    if (int rc = mg_apply(h, 0, h->g_p.p, h->g_q.p, vals0, mixed)) return rc;
    pa::k_dot<<<G, 256, 0, s>>>(h->g_p.p, h->g_q.p, ndof, PS(pa::kPPQ));
    PA_HIP(pa::launch_pcgm_alpha(PS(pa::kPPQ), G, h->scal.p, s));
    PA_HIP(pa::launch_mg_update(x, h->g_r.p, h->g_p.p, h->g_q.p, h->scal.p, nn, PS(pa::kPRR), G, s));
    if (int rc = precondition(pa::kBetaStep)) return rc;
    ++it;
    const bool last = it >= opts.max_iters;
    bool restarted = false;
    if (it % check_every == 0 || last) {
      if (int rc = read_scal(h)) return rc;
      const double rr_it = h->host_scal[pa::kRR];
      if (mg_trace())
        std::fprintf(stderr, "mg pcg it %d: iterated rel %.3e, rz %.3e, alpha %.3e, beta %.3e\n", it,
                     std::sqrt(rr_it / bb), h->host_scal[pa::kRZ], h->host_scal[pa::kAlpha], h->host_scal[pa::kBeta]);
      if (!std::isfinite(rr_it)) return fail(PA_ERR_NOT_CONVERGED, "multigrid PCG: non-finite residual");
      // r.V(r) <= 0: the V-cycle is indefinite (a smoother interval misses the top of its
      // level's spectrum).  Widen every interval and restart PCG from the true residual.
      const bool indefinite = !(h->host_scal[pa::kRZ] > 0.0) && rr_it > opts.tolerance * opts.tolerance * bb;
      if (indefinite && !last && widenings < kMgMaxWidenings) {
        ++widenings;
        for (size_t l = 0; l + 1 < h->mg.size(); ++l) mg_set_interval(h, *h->mg[l], kMgWiden * h->mg[l]->lmax);
        if (mixed) {
          ++replacements;
          if (int rc = true_residual(&rr_ref)) return rc;
          rel = std::sqrt(rr_ref / bb);
        } else {
          pa::k_spmv<<<pa::red_grid((size_t)nn, 16), 256, 0, s>>>(vals0, h->rowmap.p, h->colidx.p, x, b, h->g_r.p,
                                                                  nn, 1);
          double rr = 0.0;
          if (int rc = mg_dot(h, h->g_r.p, h->g_r.p, ndof, pa::kDot0, &rr)) return rc;
          rel = std::sqrt(rr / bb);
        }
        if (mg_trace()) std::fprintf(stderr, "mg pcg it %d: indefinite V-cycle, intervals widened, true rel %.3e\n", it, rel);
        if (!std::isfinite(rel)) return fail(PA_ERR_NOT_CONVERGED, "multigrid PCG: non-finite residual");
        conv = rel <= opts.tolerance;
        if (!conv) {
          if (int rc = precondition(pa::kBetaInit)) return rc;
          restarted = true;
        }
      } else if (!mixed) {
        rel = std::sqrt(rr_it / bb);
        conv = rel <= opts.tolerance;
      } else if (rr_it <= delta * delta * rr_ref || rr_it <= opts.tolerance * opts.tolerance * bb || last) {
        // replace the iterated residual by the true FP64 one
        ++replacements;
        if (int rc = true_residual(&rr_ref)) return rc;
        rel = std::sqrt(rr_ref / bb);
        if (mg_trace()) std::fprintf(stderr, "mg pcg it %d: true rel %.3e\n", it, rel);
        if (!std::isfinite(rel)) return fail(PA_ERR_NOT_CONVERGED, "multigrid PCG: non-finite residual");
        conv = rel <= opts.tolerance;
        if (!conv && !last) {
          // keep the direction unless the true residual is well above the iterated one
          // (then the kept direction no longer matches r and PCG restarts from p = z)
          restarted = rr_ref > 4.0 * rr_it;
          if (int rc = precondition(restarted ? pa::kBetaInit : pa::kBetaReplace)) return rc;
        }
      }
    }
    if (!conv && !last && !restarted) PA_HIP(pa::launch_mg_pupdate(h->g_p.p, h->g_z.p, h->scal.p, ndof, s));
  }
  double true_rel = rel;
  if (!mixed) {  // true FP64 residual of the assembled system
    pa::k_spmv<<<pa::red_grid((size_t)nn, 16), 256, 0, s>>>(vals0, h->rowmap.p, h->colidx.p, x, b, h->g_q.p, nn, 1);
    double rr = 0.0;
    if (int rc = mg_dot(h, h->g_q.p, h->g_q.p, ndof, pa::kDot0, &rr)) return rc;
    true_rel = bb > 0 ? std::sqrt(rr / bb) : 0.0;
  }
  if (st) {
    st->iterations = it;
    st->outer_iterations = mixed ? replacements : 1;
    st->converged = conv ? 1 : 0;
    st->rel_residual = true_rel;
  }
  if (!conv)
    return fail(PA_ERR_NOT_CONVERGED, "multigrid PCG stopped at relative residual " + std::to_string(rel));
  return PA_OK;
}

int pa_smooth(pa_handle h, const double* z, double* u, int steps, double omega) {
  if (int rc = ready(h)) return rc;
  if (!z || !u) return fail(PA_ERR_ARG, "pa_smooth: null pointer");
  if (steps < 0 || !(omega > 0.0 && omega < 2.0)) return fail(PA_ERR_ARG, "pa_smooth: steps >= 0, 0 < omega < 2");
  if (int rc = mixed_setup(h, z)) return rc;
  for (int k = 0; k < steps; ++k) {
    if (int rc = residual(h, u, z, h->m_R.p, false)) return rc;  // K u
    PA_HIP(pa::launch_jacobi_step(u, h->m_R.p, h->m_dinv.p, h->bcmask.p, h->nn, omega, h->stream));
  }
  return PA_OK;
}

int pa_solution(pa_handle h, const double* z, double* u, double* vals, const pa_solver_opts* o,
                pa_solver_stats* st) {
  if (int rc = ready(h)) return rc;
  if (!z || !u) return fail(PA_ERR_ARG, "pa_solution: null pointer");
  pa_solver_opts opts = o ? *o : pa_solver_opts{1e-10, 100000, 16, 0, 0, 0.1};
  if (opts.check_every <= 0) opts.check_every = 16;
  if (opts.preconditioner != 0 && opts.preconditioner != 1)
    return fail(PA_ERR_ARG, "preconditioner must be 0 (block Jacobi) or 1 (multigrid)");
  const bool mg = opts.preconditioner == 1;
  if (opts.precision == 1)
    return mg ? solve_mg(h, z, u, nullptr, nullptr, opts, st, true) : solve_mixed(h, z, u, nullptr, opts, st);
  if (opts.precision != 0) return fail(PA_ERR_ARG, "precision must be 0 (FP64) or 1 (mixed)");
  if (!vals) return fail(PA_ERR_ARG, "pa_solution: FP64 solve needs the values buffer");
  if (int rc = ensure_solver_ws(h)) return rc;
  const size_t ndof = 3 * (size_t)h->nn;
  // rhs = -R(0,z) = f
  double* rhs = h->w_q.p;  // reused: pa_solve copies b before touching w_q
  if (!h->w_vals.p || h->w_vals.n != ndof) PA_HIP(h->w_vals.alloc(ndof));
  rhs = h->w_vals.p;
  pa::k_fill<<<grid_for(ndof), 256, 0, h->stream>>>(rhs, ndof, 0.0);
  if (int rc = sub_load(h, rhs, 1.0)) return rc;
  if (int rc = assemble_jacobian(h, nullptr, z, vals, nullptr)) return rc;
  if (int rc = pa_apply_constraints(h, vals, rhs, 1.0)) return rc;
  if (opts.use_x0) {  // warm start: the caller's u, its constrained dofs at their prescribed values
    const int nbc = (int)h->bc_dofs.size();
    if (nbc) pa::k_bc_set<<<grid_for(nbc), 256, 0, h->stream>>>(u, h->bc_d.p, h->bc_v.p, nbc, 1.0);
    PA_HIP(hipGetLastError());
  }
  if (mg) return solve_mg(h, z, u, rhs, vals, opts, st, false);
  return pa_solve(h, vals, u, rhs, &opts, st);
}

// Problem::solution (FP64, assembled K, block Jacobi) by deflated PCG over the k vectors W.
int pa_solution_deflated(pa_handle h, const double* z, double* u, double* vals, const double* Wv, int k,
                         const pa_solver_opts* o, pa_solver_stats* st) {
  if (int rc = ready(h)) return rc;
  if (!z || !u || !vals || (k > 0 && !Wv)) return fail(PA_ERR_ARG, "pa_solution_deflated: null pointer");
  if (k < 0 || k > pa::kMaxDeflation)
    return fail(PA_ERR_ARG, "pa_solution_deflated: k must be 0.." + std::to_string(pa::kMaxDeflation));
  pa_solver_opts opts;
  if (int rc = solve_opts(o, opts, "pa_solution_deflated")) return rc;
  if (opts.precision != 0) return fail(PA_ERR_ARG, "pa_solution_deflated: FP64 (precision 0) only");
  if (int rc = ensure_solver_ws(h)) return rc;
  const int nn = h->nn;
  const size_t ndof = 3 * (size_t)nn;
  hipStream_t s = h->stream;
  if (!h->w_vals.p || h->w_vals.n != ndof) PA_HIP(h->w_vals.alloc(ndof));
  double* rhs = h->w_vals.p;
  pa::k_fill<<<grid_for(ndof), 256, 0, s>>>(rhs, ndof, 0.0);
  if (int rc = sub_load(h, rhs, 1.0)) return rc;
  if (int rc = assemble_jacobian(h, nullptr, z, vals, nullptr)) return rc;
  if (int rc = pa_apply_constraints(h, vals, rhs, 1.0)) return rc;
  const int nbc = (int)h->bc_dofs.size();
  if (opts.use_x0 && nbc) pa::k_bc_set<<<grid_for(nbc), 256, 0, s>>>(u, h->bc_d.p, h->bc_v.p, nbc, 1.0);
  PcgWork work{h->w_r.p, h->w_z.p, h->w_p.p, h->w_q.p, h->w_dinv.p, h->partials.p, h->scal.p};
  if (k == 0) return pcg_block_jacobi(h, vals, u, rhs, opts, st, work);
  // V = W with the constrained dofs zeroed, A V, E = V^T A V
  const size_t kn = (size_t)k * ndof;
  if (h->d_V.n < kn) {
    PA_HIP(h->d_V.alloc(kn));
    PA_HIP(h->d_AV.alloc(kn));
  }
  if (h->d_P.n == 0) {
    PA_HIP(h->d_P.alloc((size_t)pa::kMaxDeflation * pa::kRedBlocks));
    PA_HIP(h->d_E.alloc((size_t)pa::kMaxDeflation * pa::kMaxDeflation));
    PA_HIP(h->d_Einv.alloc((size_t)pa::kMaxDeflation * pa::kMaxDeflation));
    PA_HIP(h->d_mu.alloc(pa::kMaxDeflation));
  }
  PA_HIP(hipMemcpyAsync(h->d_V.p, Wv, kn * sizeof(double), hipMemcpyDeviceToDevice, s));
  const int G = pa::red_grid((size_t)nn, 256);
  const int Gs = pa::red_grid((size_t)nn, 16);
  std::vector<double> I((size_t)k * k, 0.0);
  for (int j = 0; j < k; ++j) I[(size_t)j * k + j] = 1.0;
  PA_HIP(hipMemcpyAsync(h->d_Einv.p, I.data(), I.size() * sizeof(double), hipMemcpyHostToDevice, s));
  for (int j = 0; j < k; ++j) {
    double* v = h->d_V.p + (size_t)j * ndof;
    if (nbc) pa::k_bc_zero_vec<<<grid_for(nbc), 256, 0, s>>>(v, h->bc_d.p, nbc);
    pa::k_spmv<<<Gs, 256, 0, s>>>(vals, h->rowmap.p, h->colidx.p, v, nullptr, h->d_AV.p + (size_t)j * ndof, nn, 0);
  }
  for (int i = 0; i < k; ++i) {  // row i of E: (A V)^T v_i through the identity "inverse"
    pa::k_multi_dot<<<G, 256, 0, s>>>(h->d_AV.p, k, ndof, h->d_V.p + (size_t)i * ndof, nn, h->d_P.p);
    pa::k_defl_mu<<<1, 256, 0, s>>>(h->d_P.p, G, k, h->d_Einv.p, h->d_E.p + (size_t)i * k);
  }
  PA_HIP(hipGetLastError());
  std::vector<double> E((size_t)k * k);
  PA_HIP(hipMemcpyAsync(E.data(), h->d_E.p, E.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  PA_HIP(hipStreamSynchronize(s));
  // inverse of E on the vectors that are A-independent: Cholesky with diagonal pivoting,
  // a vector whose remaining pivot falls below 1e-12 of the largest diagonal is dropped
  std::vector<double> Ein((size_t)k * k, 0.0);
  {
    std::vector<double> A(E);
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < i; ++j) A[(size_t)i * k + j] = A[(size_t)j * k + i] = 0.5 * (E[(size_t)i * k + j] + E[(size_t)j * k + i]);
    double dmax = 0.0;
    for (int i = 0; i < k; ++i) dmax = std::max(dmax, A[(size_t)i * k + i]);
    std::vector<int> keep;
    std::vector<double> L;  // Cholesky factor of the kept principal submatrix, row-major m x m
    for (int c = 0; c < k; ++c) {
      // try to add column c: solve L y = A[keep, c], pivot = A_cc - y.y
      const int m = (int)keep.size();
      std::vector<double> y(m);
      for (int i = 0; i < m; ++i) {
        double t = A[(size_t)keep[i] * k + c];
        for (int j = 0; j < i; ++j) t -= L[(size_t)i * k + j] * y[j];
        y[i] = t / L[(size_t)i * k + i];
      }
      double piv = A[(size_t)c * k + c];
      for (int i = 0; i < m; ++i) piv -= y[i] * y[i];
      if (!(piv > 1e-12 * dmax)) continue;
      L.resize((size_t)k * k, 0.0);
      for (int i = 0; i < m; ++i) L[(size_t)m * k + i] = y[i];
      L[(size_t)m * k + m] = std::sqrt(piv);
      keep.push_back(c);
    }
    const int m = (int)keep.size();
    // inverse of the kept block: (L L^T)^-1 column by column
    for (int col = 0; col < m; ++col) {
      std::vector<double> e(m, 0.0), y(m), xv(m);
      e[col] = 1.0;
      for (int i = 0; i < m; ++i) {
        double t = e[i];
        for (int j = 0; j < i; ++j) t -= L[(size_t)i * k + j] * y[j];
        y[i] = t / L[(size_t)i * k + i];
      }
      for (int i = m - 1; i >= 0; --i) {
        double t = y[i];
        for (int j = i + 1; j < m; ++j) t -= L[(size_t)j * k + i] * xv[j];
        xv[i] = t / L[(size_t)i * k + i];
      }
      for (int i = 0; i < m; ++i) Ein[(size_t)keep[i] * k + keep[col]] = xv[i];
    }
  }
  PA_HIP(hipMemcpyAsync(h->d_Einv.p, Ein.data(), Ein.size() * sizeof(double), hipMemcpyHostToDevice, s));
  const Deflation D{h->d_V.p, h->d_AV.p, k, h->d_Einv.p, h->d_mu.p, h->d_P.p};
  return pcg_block_jacobi(h, vals, u, rhs, opts, st, work, &D);
}

// ---------------------------------------------------------------- stress p-norm + adjoint
static int pnorm_value(pa_handle h, const double* u, const double* z, int measure, double pn, double q, double* S) {
  pa::ChunkArgs a = h->chunk_args();
  a.u = u;
  a.z = z;
  a.measure = measure;
  a.pn = pn;
  a.q = q;
  PA_HIP(pa::launch_chunk(h->code_bits, pa::kModePNorm, a, h->stream));
  return energy_value(h, S);
}

static int check_pnorm_args(int measure, double& pn, double q) {
  if (measure != 0 && measure != 1) return fail(PA_ERR_ARG, "p-norm measure must be 0 (Voigt) or 1 (von Mises)");
  if (measure == 0) pn = std::floor(pn);  // TensorPNormFunctor takes an integer exponent
  if (!(pn >= 1.0)) return fail(PA_ERR_ARG, "p-norm exponent must be >= 1");
  if (measure == 1 && !(q >= 0.0)) return fail(PA_ERR_ARG, "relaxation exponent q must be >= 0");
  return PA_OK;
}

int pa_stress_pnorm(pa_handle h, const double* u, const double* z, int measure, double pn, double q, double* value,
                    double* dfdu, double* dfdz) {
  if (int rc = ready(h)) return rc;
  if (!u || !z) return fail(PA_ERR_ARG, "pa_stress_pnorm: null pointer");
  if (int rc = check_pnorm_args(measure, pn, q)) return rc;
  double S = 0.0;
  if (int rc = pnorm_value(h, u, z, measure, pn, q, &S)) return rc;
  if (value) *value = std::pow(S, 1.0 / pn);
  if (!dfdu && !dfdz) return PA_OK;
  if (int rc = ensure_solver_ws(h)) return rc;
  pa::ChunkArgs a = h->chunk_args();
  a.u = u;
  a.z = z;
  a.measure = measure;
  a.pn = pn;
  a.q = q;
  a.pscale = S > 0.0 ? std::pow(S, (1.0 - pn) / pn) / pn : 0.0;
  a.out_vec = dfdu ? dfdu : h->w_r.p;
  a.out_vec2 = dfdz ? dfdz : h->w_z.p;  // w_z holds >= num_nodes doubles
  PA_HIP(pa::launch_chunk(h->code_bits, pa::kModePNormGrad, a, h->stream));
  return PA_OK;
}

int pa_stress_pnorm_gradient_x(pa_handle h, const double* u, const double* z, int measure, double pn, double q,
                               double* dfdx, double* value) {
  if (int rc = ready(h)) return rc;
  if (!u || !z || !dfdx) return fail(PA_ERR_ARG, "pa_stress_pnorm_gradient_x: null pointer");
  if (int rc = check_pnorm_args(measure, pn, q)) return rc;
  double S = 0.0;
  if (int rc = pnorm_value(h, u, z, measure, pn, q, &S)) return rc;
  if (value) *value = std::pow(S, 1.0 / pn);
  pa::ChunkArgs a = h->chunk_args();
  a.u = u;
  a.z = z;
  a.measure = measure;
  a.pn = pn;
  a.q = q;
  a.pscale = S > 0.0 ? std::pow(S, (1.0 - pn) / pn) / pn : 0.0;
  a.out_vec = dfdx;
  PA_HIP(pa::launch_chunk(h->code_bits, pa::kModePNormGradX, a, h->stream));
  return PA_OK;
}

int pa_dRdz_T_apply(pa_handle h, const double* u, const double* lambda, const double* z, double* out) {
  if (int rc = ready(h)) return rc;
  if (!u || !lambda || !z || !out) return fail(PA_ERR_ARG, "pa_dRdz_T_apply: null pointer");
  pa::ChunkArgs a = h->chunk_args();
  a.u = u;
  a.u2 = lambda;
  a.z = z;
  a.out_vec = out;
  PA_HIP(pa::launch_chunk(h->code_bits, pa::kModeAdjZ, a, h->stream));
  return PA_OK;
}

int pa_matvec_plus_vector(pa_handle h, int block_rows, double alpha, const double* values, const double* x,
                          double* y) {
  if (int rc = ready(h)) return rc;
  if (!values || !x || !y) return fail(PA_ERR_ARG, "pa_matvec_plus_vector: null pointer");
  if (block_rows != 1 && block_rows != 3) return fail(PA_ERR_ARG, "pa_matvec_plus_vector: block_rows must be 1 or 3");
  const int Gs = pa::red_grid((size_t)h->nn, 16);
  if (block_rows == 3)
    pa::k_matvec_add<3><<<Gs, 256, 0, h->stream>>>(values, h->rowmap.p, h->colidx.p, x, y, h->nn, alpha);
  else
    pa::k_matvec_add<1><<<Gs, 256, 0, h->stream>>>(values, h->rowmap.p, h->colidx.p, x, y, h->nn, alpha);
  PA_HIP(hipGetLastError());
  return PA_OK;
}

// Total derivative of the p-norm through K(z) u = f (Problem::criterionGradient, non
// self-adjoint branch, src/elliptic/Problem.hpp:413-468): K lambda = dF/du with
// homogeneous constraints, dF/dz = dF/dz|_u - (dR/dz)^T lambda.  lambda (3N, device)
// receives the adjoint state.
int pa_stress_pnorm_gradient(pa_handle h, const double* u, const double* z, int measure, double pn, double q,
                             double* dfdz, double* lambda, double* value, const pa_solver_opts* o,
                             pa_solver_stats* st) {
  if (int rc = ready(h)) return rc;
  if (!u || !z || !dfdz || !lambda) return fail(PA_ERR_ARG, "pa_stress_pnorm_gradient: null pointer");
  pa_solver_opts opts = o ? *o : pa_solver_opts{1e-10, 100000, 16, 0, 0, 0.1};
  if (opts.check_every <= 0) opts.check_every = 16;
  if (int rc = ensure_solver_ws(h)) return rc;
  const size_t ndof = 3 * (size_t)h->nn;
  if (h->w_adj.n != ndof) PA_HIP(h->w_adj.alloc(ndof));
  double* dfdu = h->w_adj.p;
  if (int rc = pa_stress_pnorm(h, u, z, measure, pn, q, value, dfdu, dfdz)) return rc;
  opts.use_x0 = 0;
  if (opts.preconditioner != 0 && opts.preconditioner != 1)
    return fail(PA_ERR_ARG, "preconditioner must be 0 (block Jacobi) or 1 (multigrid)");
  const bool mg = opts.preconditioner == 1;
  if (opts.precision == 1) {
    if (int rc = mg ? solve_mg(h, z, lambda, dfdu, nullptr, opts, st, true) : solve_mixed(h, z, lambda, dfdu, opts, st))
      return rc;
  } else {
    const size_t nv = 9 * (size_t)h->graph.rowmap[h->nn];
    if (h->w_K.n != nv) PA_HIP(h->w_K.alloc(nv));
    if (int rc = assemble_jacobian(h, nullptr, z, h->w_K.p, nullptr)) return rc;
    if (int rc = pa_apply_constraints(h, h->w_K.p, dfdu, 0.0)) return rc;  // homogeneous on lambda
    if (int rc = mg ? solve_mg(h, z, lambda, dfdu, h->w_K.p, opts, st, false) : pa_solve(h, h->w_K.p, lambda, dfdu, &opts, st))
      return rc;
  }
  if (h->w_adjz.n != (size_t)h->nn) PA_HIP(h->w_adjz.alloc(h->nn));
  if (int rc = pa_dRdz_T_apply(h, u, lambda, z, h->w_adjz.p)) return rc;
  pa::k_axpy<<<grid_for((size_t)h->nn), 256, 0, h->stream>>>(dfdz, h->w_adjz.p, (size_t)h->nn, -1.0);
  PA_HIP(hipGetLastError());
  return PA_OK;
}

int pa_compliance_gradient(pa_handle h, const double* u, const double* z, double* dcdz, double* value) {
  if (int rc = pa_sf_gradient_z(h, u, z, dcdz, value)) return rc;
  pa::k_scale<<<grid_for((size_t)h->nn), 256, 0, h->stream>>>(dcdz, (size_t)h->nn, -1.0);
  PA_HIP(hipGetLastError());
  return PA_OK;
}

}  // extern "C"
