// Block-CRS (3x3) kernels: SpMV, Jacobi-PCG steps with fused grid reductions,
// Dirichlet elimination and load application.
#include "pa_solver.hpp"

namespace pa {

constexpr int kT = 256;
constexpr int kLanesPerRow = 16;  // 4 block rows per wavefront

// ------------------------------------------------------------------ reductions
// Block sum of one value per thread (deterministic order).
__device__ __forceinline__ double block_sum(double v, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int q = 0; q < kT / 64; ++q) t += sh[q];
  return t;  // valid in thread 0
}

// Returns true in every thread of the block that arrives last.  Partials written
// before the call are visible to that block.  The counter is reset by the last block.
__device__ __forceinline__ bool last_block(unsigned* counter) {
  __shared__ int am_last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned t = atomicAdd(counter, 1u);
    am_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (am_last) __threadfence();
  return am_last != 0;
}

__device__ __forceinline__ double sum_partials(const double* p, int n, double* sh) {
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += kT) v += __builtin_nontemporal_load(p + i);
  return block_sum(v, sh);
}

// ------------------------------------------------------------------ SpMV
// 16 lanes per block row walk its 9*len values contiguously; lane partials for the
// three output components are reduced with cross-lane shuffles.
__device__ __forceinline__ void spmv_row(const double* __restrict__ vals, const int* __restrict__ rowmap,
                                         const int* __restrict__ col, const double* __restrict__ x, int row,
                                         int lane, double acc[3]) {
  acc[0] = acc[1] = acc[2] = 0.0;
  const int s0 = rowmap[row], s1 = rowmap[row + 1];
  const int n = 9 * (s1 - s0);
  const double* v = vals + 9 * (size_t)s0;
  int k = lane;
  int s = k / 9, t = k - 9 * s;
  for (; k < n; k += kLanesPerRow) {
    const int c = col[s0 + s];
    const int i = t / 3, j = t - 3 * (t / 3);
    const double p = v[k] * x[3 * (size_t)c + j];
    if (i == 0) acc[0] += p;
    else if (i == 1) acc[1] += p;
    else acc[2] += p;
    t += kLanesPerRow - 9;  // advance 16 = 9 + 7
    s += 1;
    if (t >= 9) {
      t -= 9;
      s += 1;
    }
  }
#pragma unroll
  for (int off = kLanesPerRow / 2; off > 0; off >>= 1) {
    acc[0] += __shfl_xor(acc[0], off, kLanesPerRow);
    acc[1] += __shfl_xor(acc[1], off, kLanesPerRow);
    acc[2] += __shfl_xor(acc[2], off, kLanesPerRow);
  }
}

// y = A x (mode 0) or y = b - A x (mode 1)
__global__ __launch_bounds__(kT) void k_spmv(const double* vals, const int* rowmap, const int* col,
                                             const double* x, const double* b, double* y, int nrows,
                                             int mode) {
  const int lane = threadIdx.x % kLanesPerRow;
  const int rows_per_block = kT / kLanesPerRow;
  for (int row = blockIdx.x * rows_per_block + threadIdx.x / kLanesPerRow; row < nrows;
       row += gridDim.x * rows_per_block) {
    double acc[3];
    spmv_row(vals, rowmap, col, x, row, lane, acc);
    if (lane < 3) {
      const double a = lane == 0 ? acc[0] : (lane == 1 ? acc[1] : acc[2]);
      const size_t d = 3 * (size_t)row + lane;
      y[d] = mode == 0 ? a : b[d] - a;
    }
  }
}

// q = A p, then alpha = rz / (p.q) by the last block.
__global__ __launch_bounds__(kT) void k_pcg_spmv(const double* vals, const int* rowmap, const int* col,
                                                 const double* p, double* q, int nrows, double* partials,
                                                 unsigned* counter, double* scal) {
  __shared__ double sh[kT / 64];
  const int lane = threadIdx.x % kLanesPerRow;
  const int rows_per_block = kT / kLanesPerRow;
  double dot = 0.0;
  for (int row = blockIdx.x * rows_per_block + threadIdx.x / kLanesPerRow; row < nrows;
       row += gridDim.x * rows_per_block) {
    double acc[3];
    spmv_row(vals, rowmap, col, p, row, lane, acc);
    if (lane < 3) {
      const double a = lane == 0 ? acc[0] : (lane == 1 ? acc[1] : acc[2]);
      const size_t d = 3 * (size_t)row + lane;
      q[d] = a;
      dot += a * p[d];
    }
  }
  const double t = block_sum(dot, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
  if (last_block(counter)) {
    const double pq = sum_partials(partials, gridDim.x, sh);
    if (threadIdx.x == 0) {
      scal[kPQ] = pq;
      scal[kAlpha] = scal[kRZ] / pq;
      *counter = 0;
    }
  }
}

// z = Dinv r on 3x3 blocks
__device__ __forceinline__ void apply_dinv(const double* dinv, const double* r, double* z, int n) {
  const double* D = dinv + 9 * (size_t)n;
  const double r0 = r[3 * (size_t)n], r1 = r[3 * (size_t)n + 1], r2 = r[3 * (size_t)n + 2];
  z[3 * (size_t)n + 0] = D[0] * r0 + D[1] * r1 + D[2] * r2;
  z[3 * (size_t)n + 1] = D[3] * r0 + D[4] * r1 + D[5] * r2;
  z[3 * (size_t)n + 2] = D[6] * r0 + D[7] * r1 + D[8] * r2;
}

// x += alpha p; r -= alpha q; z = Dinv r; rz_new = r.z, rr = r.r; beta by the last block.
__global__ __launch_bounds__(kT) void k_pcg_update(double* x, double* r, double* z, const double* p,
                                                   const double* q, const double* dinv, int nnodes,
                                                   double* partials, unsigned* counter, double* scal) {
  __shared__ double sh[kT / 64];
  const double alpha = scal[kAlpha];
  double rz = 0.0, rr = 0.0;
  for (int n = blockIdx.x * kT + threadIdx.x; n < nnodes; n += gridDim.x * kT) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const size_t d = 3 * (size_t)n + i;
      x[d] += alpha * p[d];
      r[d] -= alpha * q[d];
    }
    apply_dinv(dinv, r, z, n);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const size_t d = 3 * (size_t)n + i;
      rz += r[d] * z[d];
      rr += r[d] * r[d];
    }
  }
  const double t0 = block_sum(rz, sh);
  const double t1 = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = t0;
    partials[2 * blockIdx.x + 1] = t1;
  }
  if (last_block(counter)) {
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += kT) {
      a += partials[2 * i];
      b += partials[2 * i + 1];
    }
    const double sa = block_sum(a, sh);
    const double sb = block_sum(b, sh);
    if (threadIdx.x == 0) {
      scal[kBeta] = sa / scal[kRZ];
      scal[kRZ] = sa;
      scal[kRR] = sb;
      *counter = 0;
    }
  }
}

__global__ __launch_bounds__(kT) void k_pcg_pupdate(double* p, const double* z, int ndof, const double* scal) {
  const double beta = scal[kBeta];
  for (int d = blockIdx.x * kT + threadIdx.x; d < ndof; d += gridDim.x * kT) p[d] = z[d] + beta * p[d];
}

// z = Dinv r, p = z, rz = r.z, rr = r.r, bb = b.b (start of PCG)
__global__ __launch_bounds__(kT) void k_pcg_init(const double* r, double* z, double* p, const double* b,
                                                 const double* dinv, int nnodes, double* partials,
                                                 unsigned* counter, double* scal) {
  __shared__ double sh[kT / 64];
  double rz = 0.0, rr = 0.0, bb = 0.0;
  for (int n = blockIdx.x * kT + threadIdx.x; n < nnodes; n += gridDim.x * kT) {
    apply_dinv(dinv, r, z, n);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const size_t d = 3 * (size_t)n + i;
      p[d] = z[d];
      rz += r[d] * z[d];
      rr += r[d] * r[d];
      bb += b[d] * b[d];
    }
  }
  const double t0 = block_sum(rz, sh);
  const double t1 = block_sum(rr, sh);
  const double t2 = block_sum(bb, sh);
  if (threadIdx.x == 0) {
    partials[3 * blockIdx.x] = t0;
    partials[3 * blockIdx.x + 1] = t1;
    partials[3 * blockIdx.x + 2] = t2;
  }
  if (last_block(counter)) {
    double a = 0.0, b2 = 0.0, c = 0.0;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += kT) {
      a += partials[3 * i];
      b2 += partials[3 * i + 1];
      c += partials[3 * i + 2];
    }
    const double sa = block_sum(a, sh);
    const double sb = block_sum(b2, sh);
    const double sc = block_sum(c, sh);
    if (threadIdx.x == 0) {
      scal[kRZ] = sa;
      scal[kRR] = sb;
      scal[kBB] = sc;
      *counter = 0;
    }
  }
}

// generic dot products: scal[slot] = a.b  (and a.a into slot+1 when want_aa)
__global__ __launch_bounds__(kT) void k_dot(const double* a, const double* b, size_t n, double* partials,
                                            unsigned* counter, double* scal, int slot) {
  __shared__ double sh[kT / 64];
  double v = 0.0;
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) v += a[i] * b[i];
  const double t = block_sum(v, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
  if (last_block(counter)) {
    const double s = sum_partials(partials, gridDim.x, sh);
    if (threadIdx.x == 0) {
      scal[slot] = s;
      *counter = 0;
    }
  }
}

// sum of per-chunk partials into scal[slot]
__global__ __launch_bounds__(kT) void k_sum(const double* partials, int n, double* scal, int slot) {
  __shared__ double sh[kT / 64];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += kT) v += partials[i];
  const double t = block_sum(v, sh);
  if (threadIdx.x == 0) scal[slot] = t;
}

// 3x3 block-diagonal inverse (cofactors)
__global__ __launch_bounds__(kT) void k_diag_inv(const double* vals, const int* diag_slot, double* dinv,
                                                 int nnodes) {
  const int n = blockIdx.x * kT + threadIdx.x;
  if (n >= nnodes) return;
  const double* a = vals + 9 * (size_t)diag_slot[n];
  const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
  const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
  const double id = 1.0 / det;
  double* o = dinv + 9 * (size_t)n;
  o[0] = c00 * id;
  o[1] = (a[2] * a[7] - a[1] * a[8]) * id;
  o[2] = (a[1] * a[5] - a[2] * a[4]) * id;
  o[3] = c01 * id;
  o[4] = (a[0] * a[8] - a[2] * a[6]) * id;
  o[5] = (a[2] * a[3] - a[0] * a[5]) * id;
  o[6] = c02 * id;
  o[7] = (a[1] * a[6] - a[0] * a[7]) * id;
  o[8] = (a[0] * a[4] - a[1] * a[3]) * id;
}

// ------------------------------------------------------------------ loads and BCs
// y[3n+i] += s * f[3k+i] for the loaded nodes (sparse traction vector).
__global__ __launch_bounds__(kT) void k_add_load(double* y, const int* nodes, const double* f, int n, double s) {
  const int k = blockIdx.x * kT + threadIdx.x;
  if (k >= n) return;
  const size_t d = 3 * (size_t)nodes[k];
  y[d] += s * f[3 * k];
  y[d + 1] += s * f[3 * k + 1];
  y[d + 2] += s * f[3 * k + 2];
}

__global__ __launch_bounds__(kT) void k_fill(double* y, size_t n, double v) {
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) y[i] = v;
}

__global__ __launch_bounds__(kT) void k_axpy(double* y, const double* x, size_t n, double a) {
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) y[i] += a * x[i];
}

__global__ __launch_bounds__(kT) void k_scale(double* y, size_t n, double s) {
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) y[i] *= s;
}

// applyBlockConstraints first loop, rhs part: rhs[col] -= A(dof,col) * value for the
// BC rows (runs before any entry is zeroed, so each lift sees the assembled matrix).
__global__ __launch_bounds__(kT) void k_bc_lift(const double* vals, const int* rowmap, const int* col,
                                                const int* dofs, const double* bcv, int nbc, double scale,
                                                double* rhs) {
  const int k = blockIdx.x * kT + threadIdx.x;
  if (k >= nbc) return;
  const double v = scale * bcv[k];
  if (v == 0.0) return;
  const int d = dofs[k], n = d / 3, i = d % 3;
  for (int s = rowmap[n]; s < rowmap[n + 1]; ++s) {
    const int c = col[s];
    for (int j = 0; j < 3; ++j) {
      const int cd = 3 * c + j;
      if (cd == d) continue;
      atomicAdd(&rhs[cd], -vals[9 * (size_t)s + 3 * i + j] * v);
    }
  }
}

// applyBlockConstraints first loop, matrix part: BC row -> unit diagonal, zero
// elsewhere; the symmetric (col,row) entries zeroed too.
__global__ __launch_bounds__(kT) void k_bc_zero(double* vals, const int* rowmap, const int* col,
                                                const int* dofs, int nbc) {
  const int k = blockIdx.x * kT + threadIdx.x;
  if (k >= nbc) return;
  const int d = dofs[k], n = d / 3, i = d % 3;
  for (int s = rowmap[n]; s < rowmap[n + 1]; ++s) {
    const int c = col[s];
    for (int j = 0; j < 3; ++j) {
      const int cd = 3 * c + j;
      vals[9 * (size_t)s + 3 * i + j] = (cd == d) ? 1.0 : 0.0;
      if (cd == d) continue;
      for (int t = rowmap[c]; t < rowmap[c + 1]; ++t)
        if (col[t] == n) vals[9 * (size_t)t + 3 * j + i] = 0.0;
    }
  }
}

// second loop: rhs(dof) = scale * value
__global__ __launch_bounds__(kT) void k_bc_set(double* rhs, const int* dofs, const double* bcv, int nbc,
                                               double scale) {
  const int k = blockIdx.x * kT + threadIdx.x;
  if (k >= nbc) return;
  rhs[dofs[k]] = scale * bcv[k];
}

// zero the BC dofs of a vector (homogeneous constraint on a correction)
__global__ __launch_bounds__(kT) void k_bc_zero_vec(double* y, const int* dofs, int nbc) {
  const int k = blockIdx.x * kT + threadIdx.x;
  if (k >= nbc) return;
  y[dofs[k]] = 0.0;
}

}  // namespace pa
