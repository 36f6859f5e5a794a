// C ABI of the Tet4 SIMP elastostatics library (declared in include/plato_mi355x.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/plato_mi355x.h"
#include "pa_host.hpp"
#include "pa_kernels.hpp"
#include "pa_mf.hpp"
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

struct pa_problem {
  int nn = 0, ne = 0;
  std::vector<double> coords;
  std::vector<int32_t> conn;
  double E = 1.0, nu = 0.3;
  double p = 3.0, rmin = 1e-9;
  std::vector<int32_t> bc_dofs;
  std::vector<double> bc_vals;
  std::vector<double> f;  // dense load, filled by pa_add_traction
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
  // solver workspace (lazily allocated)
  DevArray<double> w_r, w_z, w_p, w_q, w_dinv, w_vals, w_adj, w_adjz, w_K;
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
  if (!h) return fail(PA_ERR_ARG, "null handle");
  if (!std::isfinite(exponent) || !std::isfinite(minimum_value)) return fail(PA_ERR_ARG, "non-finite penalty");
  h->p = exponent;
  h->rmin = minimum_value;
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
  for (int mode = 0; mode <= pa::kModeAdjZ; ++mode)
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
  PA_HIP(hipHostMalloc(&h->host_scal, pa::kNumScal * sizeof(double), hipHostMallocDefault));
  h->finalized = true;
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

static int sub_load(pa_handle h, double* R, double s) {
  if (h->nload == 0) return PA_OK;
  pa::k_add_load<<<grid_for(h->nload), 256, 0, h->stream>>>(R, h->load_nodes.p, h->load_f.p, h->nload, s);
  PA_HIP(hipGetLastError());
  return PA_OK;
}

int pa_vf_value(pa_handle h, const double* u, const double* z, double* R) {
  if (int rc = ready(h)) return rc;
  if (!u || !z || !R) return fail(PA_ERR_ARG, "pa_vf_value: null pointer");
  return run_chunk(h, pa::kModeResidual, u, z, nullptr, R, true);
}

int pa_vf_gradient_u(pa_handle h, const double* u, const double* z, double* vals, double* R) {
  if (int rc = ready(h)) return rc;
  if (!z || !vals || (R && !u)) return fail(PA_ERR_ARG, "pa_vf_gradient_u: null pointer");
  if (R) return run_chunk(h, pa::kModeJacobianResidual, u, z, vals, R, true);
  return run_chunk(h, pa::kModeJacobian, u, z, vals, nullptr);
}

int pa_vf_gradient_z(pa_handle h, const double* u, const double* z, double* vals_t) {
  if (int rc = ready(h)) return rc;
  if (!u || !z || !vals_t) return fail(PA_ERR_ARG, "pa_vf_gradient_z: null pointer");
  return run_chunk(h, pa::kModeGradZT, u, z, vals_t, nullptr);
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
  return run_chunk(h, pa::kModeResidual, u, z, nullptr, g);
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

static int read_scal(pa_handle h) {
  PA_HIP(hipMemcpyAsync(h->host_scal, h->scal.p, pa::kNumScal * sizeof(double), hipMemcpyDeviceToHost,
                        h->stream));
  PA_HIP(hipStreamSynchronize(h->stream));
  return PA_OK;
}

int pa_solve(pa_handle h, const double* vals, double* x, const double* b, const pa_solver_opts* o,
             pa_solver_stats* st) {
  if (int rc = ready(h)) return rc;
  if (!vals || !x || !b) return fail(PA_ERR_ARG, "pa_solve: null pointer");
  pa_solver_opts opts = o ? *o : pa_solver_opts{1e-10, 100000, 16, 0, 0, 0.1};
  if (opts.check_every <= 0) opts.check_every = 16;
  if (int rc = ensure_solver_ws(h)) return rc;
  const int nn = h->nn;
  const size_t ndof = 3 * (size_t)nn;
  hipStream_t s = h->stream;
  const int G = pa::red_grid((size_t)nn, 256);       // node-parallel kernels
  const int Gs = pa::red_grid((size_t)nn, 16);       // row-group kernels (16 rows per group)
  double* P = h->partials.p;
  auto PS = [&](int slot) { return P + (size_t)slot * pa::kRedBlocks; };
  pa::k_diag_inv<<<grid_for(nn), 256, 0, s>>>(vals, h->diag_slot.p, h->w_dinv.p, nn);
  if (opts.use_x0) {
    pa::k_spmv<<<Gs, 256, 0, s>>>(vals, h->rowmap.p, h->colidx.p, x, b, h->w_r.p, nn, 1);
  } else {
    pa::k_fill<<<grid_for(ndof), 256, 0, s>>>(x, ndof, 0.0);
    PA_HIP(hipMemcpyAsync(h->w_r.p, b, ndof * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  pa::k_pcg_init<<<G, 256, 0, s>>>(h->w_r.p, h->w_z.p, h->w_p.p, b, h->w_dinv.p, nn, PS(pa::kPRZ0), PS(pa::kPRR),
                                    PS(pa::kPBB));
  pa::k_sum<<<1, 256, 0, s>>>(PS(pa::kPRR), G, h->scal.p, pa::kRR);
  pa::k_sum<<<1, 256, 0, s>>>(PS(pa::kPBB), G, h->scal.p, pa::kBB);
  PA_HIP(hipGetLastError());
  if (int rc = read_scal(h)) return rc;
  const double bb = h->host_scal[pa::kBB];
  double rel = bb > 0 ? std::sqrt(h->host_scal[pa::kRR] / bb) : 0.0;
  int it = 0;
  bool conv = bb == 0.0 || rel <= opts.tolerance;
  while (!conv && it < opts.max_iters) {
    // rz of this iteration in P_RZ[cur]; the update writes the next one into P_RZ[cur ^ 1]
    const int cur = it & 1;
    pa::k_pcg_spmv<<<Gs, 256, 0, s>>>(vals, h->rowmap.p, h->colidx.p, h->w_p.p, h->w_q.p, nn, PS(pa::kPPQ));
    pa::k_pcg_update<<<G, 256, 0, s>>>(x, h->w_r.p, h->w_z.p, h->w_p.p, h->w_q.p, h->w_dinv.p, nn, PS(pa::kPPQ),
                                       Gs, PS(pa::kPRZ0 + cur), G, PS(pa::kPRZ0 + (cur ^ 1)), PS(pa::kPRR),
                                       h->scal.p);
    pa::k_pcg_pupdate<<<G, 256, 0, s>>>(h->w_p.p, h->w_z.p, nn, PS(pa::kPRZ0 + (cur ^ 1)), PS(pa::kPRZ0 + cur), G);
    ++it;
    if (it % opts.check_every == 0 || it == opts.max_iters) {
      pa::k_sum<<<1, 256, 0, s>>>(PS(pa::kPRR), G, h->scal.p, pa::kRR);
      PA_HIP(hipGetLastError());
      if (int rc = read_scal(h)) return rc;
      rel = std::sqrt(h->host_scal[pa::kRR] / bb);
      if (!std::isfinite(rel)) return fail(PA_ERR_NOT_CONVERGED, "PCG produced a non-finite residual");
      conv = rel <= opts.tolerance;
    }
  }
  // true residual
  pa::k_spmv<<<Gs, 256, 0, s>>>(vals, h->rowmap.p, h->colidx.p, x, b, h->w_r.p, nn, 1);
  pa::k_dot<<<G, 256, 0, s>>>(h->w_r.p, h->w_r.p, ndof, PS(pa::kPDot));
  pa::k_sum<<<1, 256, 0, s>>>(PS(pa::kPDot), G, h->scal.p, pa::kDot0);
  PA_HIP(hipGetLastError());
  if (int rc = read_scal(h)) return rc;
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
  PA_HIP(pa::launch_mf_diag_inv(da, s));
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
    if (int rc = run_chunk(h, pa::kModeResidual, u, z, nullptr, h->m_R.p, b == nullptr)) return rc;
    return masked_residual(h->m_R.p, rr);
  };
  double rr_ref = 0.0;
  if (int rc = true_residual(&rr_ref)) return rc;
  rel = bb > 0 ? std::sqrt(rr_ref / bb) : 0.0;
  if (!std::isfinite(rel)) return fail(PA_ERR_NOT_CONVERGED, "mixed-precision solve: non-finite residual");
  conv = bb == 0.0 || rel <= opts.tolerance;
  if (!conv) {
    PA_HIP(pa::launch_pcgm_init(h->m_r.p, h->m_z.p, h->m_dinv.p, nn, PS(pa::kPRZ0), G, s));
    PA_HIP(pa::launch_pcgm_beta(PS(pa::kPRZ0), nullptr, G, h->scal.p, pa::kBetaInit, s));
  }
  int it = 0;
  while (!conv) {
    const int cur = it & 1;
    ma.first = it == 0;
    ma.p_prev = pbuf[cur ^ 1];
    ma.p_cur = pbuf[cur];
    PA_HIP(pa::launch_mf_apply(ma, s));
    PA_HIP(pa::launch_pcgm_alpha(h->chunk_partials.p, h->chunks.count, h->scal.p, s));
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
    // continue from the replaced residual: z = Dinv r, beta = (r.z) / rz of the last step
    PA_HIP(pa::launch_pcgm_init(h->m_r.p, h->m_z.p, h->m_dinv.p, nn, PS(pa::kPRZ0), G, s));
    PA_HIP(pa::launch_pcgm_beta(PS(pa::kPRZ0), nullptr, G, h->scal.p, pa::kBetaReplace, s));
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

int pa_smooth(pa_handle h, const double* z, double* u, int steps, double omega) {
  if (int rc = ready(h)) return rc;
  if (!z || !u) return fail(PA_ERR_ARG, "pa_smooth: null pointer");
  if (steps < 0 || !(omega > 0.0 && omega < 2.0)) return fail(PA_ERR_ARG, "pa_smooth: steps >= 0, 0 < omega < 2");
  if (int rc = mixed_setup(h, z)) return rc;
  for (int k = 0; k < steps; ++k) {
    if (int rc = run_chunk(h, pa::kModeResidual, u, z, nullptr, h->m_R.p, false)) return rc;  // K u
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
  if (opts.precision == 1) return solve_mixed(h, z, u, nullptr, opts, st);
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
  if (int rc = run_chunk(h, pa::kModeJacobian, nullptr, z, vals, nullptr)) return rc;
  if (int rc = pa_apply_constraints(h, vals, rhs, 1.0)) return rc;
  opts.use_x0 = 0;
  return pa_solve(h, vals, u, rhs, &opts, st);
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
  if (opts.precision == 1) {
    if (int rc = solve_mixed(h, z, lambda, dfdu, opts, st)) return rc;
  } else {
    const size_t nv = 9 * (size_t)h->graph.rowmap[h->nn];
    if (h->w_K.n != nv) PA_HIP(h->w_K.alloc(nv));
    if (int rc = run_chunk(h, pa::kModeJacobian, nullptr, z, h->w_K.p, nullptr)) return rc;
    if (int rc = pa_apply_constraints(h, h->w_K.p, dfdu, 0.0)) return rc;  // homogeneous on lambda
    if (int rc = pa_solve(h, h->w_K.p, lambda, dfdu, &opts, st)) return rc;
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
