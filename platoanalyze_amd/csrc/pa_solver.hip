// Block-CRS (3x3) kernels: SpMV, Jacobi-PCG steps with fence-free grid reductions,
// Dirichlet elimination and load application.
#include "pa_reduce.hpp"
#include "pa_solver.hpp"
#include "pa_spmv.hpp"

namespace pa {

constexpr int kT = 256;
constexpr int kRowsPerBlock = kT / kLanesPerRow;

// ------------------------------------------------------------------ SpMV
// y = A x (mode 0) or y = b - A x (mode 1)
__global__ __launch_bounds__(kT) void k_spmv(const double* __restrict__ vals, const int* __restrict__ rowmap,
                                             const int* __restrict__ col, const double* __restrict__ x,
                                             const double* __restrict__ b, double* __restrict__ y, int nrows,
                                             int mode) {
  const int lane = threadIdx.x % kLanesPerRow;
  const XcdSpan sp = xcd_span(nrows);
  auto X = [&](int c, int j) { return x[3 * (size_t)c + j]; };
  for (int row = sp.lo + sp.j * kRowsPerBlock + threadIdx.x / kLanesPerRow; row < sp.hi;
       row += sp.bx * kRowsPerBlock) {
    double acc[3];
    spmv_row(vals, rowmap, col, row, lane, X, acc);
    if (lane < 3) {
      const double a = lane == 0 ? acc[0] : (lane == 1 ? acc[1] : acc[2]);
      const size_t d = 3 * (size_t)row + lane;
      y[d] = mode == 0 ? a : b[d] - a;
    }
  }
}

// y += alpha A x on the block graph for 3x3 blocks (BR = 3: Jacobian and transposed
// dR/dX layouts, 9 values per entry) or 1x3 blocks (BR = 1: transposed dR/dz, 3 values
// per entry); 16 lanes per row as in the SpMV.
template <int BR>
__global__ __launch_bounds__(kT) void k_matvec_add(const double* __restrict__ vals, const int* __restrict__ rowmap,
                                                   const int* __restrict__ col, const double* __restrict__ x,
                                                   double* __restrict__ y, int nrows, double alpha) {
  const int lane = threadIdx.x % kLanesPerRow;
  const XcdSpan sp = xcd_span(nrows);
  for (int row = sp.lo + sp.j * kRowsPerBlock + threadIdx.x / kLanesPerRow; row < sp.hi;
       row += sp.bx * kRowsPerBlock) {
    double a[BR];
#pragma unroll
    for (int i = 0; i < BR; ++i) a[i] = 0.0;
    for (int s = rowmap[row] + lane; s < rowmap[row + 1]; s += kLanesPerRow) {
      const int c = col[s];
      const double* v = vals + 3 * BR * (size_t)s;
      const double x0 = x[3 * (size_t)c], x1 = x[3 * (size_t)c + 1], x2 = x[3 * (size_t)c + 2];
#pragma unroll
      for (int i = 0; i < BR; ++i) a[i] += v[3 * i] * x0 + v[3 * i + 1] * x1 + v[3 * i + 2] * x2;
    }
#pragma unroll
    for (int i = 0; i < BR; ++i) {
#pragma unroll
      for (int off = kLanesPerRow / 2; off > 0; off >>= 1) a[i] += __shfl_xor(a[i], off, kLanesPerRow);
    }
#pragma unroll
    for (int i = 0; i < BR; ++i)
      if (lane == i) y[BR * (size_t)row + i] += alpha * a[i];
  }
}
template __global__ void k_matvec_add<1>(const double*, const int*, const int*, const double*, double*, int, double);
template __global__ void k_matvec_add<3>(const double*, const int*, const int*, const double*, double*, int, double);

// q = A p; per-group partial of p.q into P_pq.
__global__ __launch_bounds__(kT) void k_pcg_spmv(const double* __restrict__ vals, const int* __restrict__ rowmap,
                                                 const int* __restrict__ col, const double* __restrict__ p,
                                                 double* __restrict__ q, int nrows, double* __restrict__ P_pq) {
  __shared__ double sh[kT / 64];
  const int lane = threadIdx.x % kLanesPerRow;
  const XcdSpan sp = xcd_span(nrows);
  auto X = [&](int c, int j) { return p[3 * (size_t)c + j]; };
  double dot = 0.0;
  for (int row = sp.lo + sp.j * kRowsPerBlock + threadIdx.x / kLanesPerRow; row < sp.hi;
       row += sp.bx * kRowsPerBlock) {
    double acc[3];
    spmv_row(vals, rowmap, col, row, lane, X, acc);
    if (lane < 3) {
      const double a = lane == 0 ? acc[0] : (lane == 1 ? acc[1] : acc[2]);
      const size_t d = 3 * (size_t)row + lane;
      q[d] = a;
      dot += p[d] * a;
    }
  }
  const double t = block_reduce_all<kT>(dot, sh);
  if (threadIdx.x == 0) P_pq[blockIdx.x] = t;
}

// beta = rz_new / rz_old from the partials of the last two update kernels; p = z + beta p.
__global__ __launch_bounds__(kT) void k_pcg_pupdate(double* __restrict__ p, const double* __restrict__ z,
                                                    int nnodes, const double* P_rz_new, const double* P_rz_old,
                                                    int np) {
  __shared__ double sh[kT / 64];
  const double a = sum_partials_all<kT>(P_rz_new, np, sh);
  const double b = sum_partials_all<kT>(P_rz_old, np, sh);
  const double beta = a / b;
  const XcdSpan sp = xcd_span(nnodes);
  for (int n = sp.lo + sp.j * kT + threadIdx.x; n < sp.hi; n += sp.bx * kT) {
    const size_t d = 3 * (size_t)n;
#pragma unroll
    for (int i = 0; i < 3; ++i) p[d + i] = fma(beta, p[d + i], z[d + i]);
  }
}

// z = Dinv r on one 3x3 block
__device__ __forceinline__ void dinv3(const double* __restrict__ D, const double r[3], double z[3]) {
  z[0] = D[0] * r[0] + D[1] * r[1] + D[2] * r[2];
  z[1] = D[3] * r[0] + D[4] * r[1] + D[5] * r[2];
  z[2] = D[6] * r[0] + D[7] * r[1] + D[8] * r[2];
}

// alpha = rz / pq from the partials; x += alpha p; r -= alpha q; z = Dinv r;
// partials of r.z (next rz) and r.r.  Group 0 records pq, rz, alpha in scal.
__global__ __launch_bounds__(kT) void k_pcg_update(double* __restrict__ x, double* __restrict__ r,
                                                   double* __restrict__ z, const double* __restrict__ p,
                                                   const double* __restrict__ q, const double* __restrict__ dinv,
                                                   int nnodes, const double* P_pq, int np_pq, const double* P_rz,
                                                   int np, double* __restrict__ P_rz_next,
                                                   double* __restrict__ P_rr, double* __restrict__ scal) {
  __shared__ double sh[kT / 64];
  const double pq = sum_partials_all<kT>(P_pq, np_pq, sh);
  const double rz = sum_partials_all<kT>(P_rz, np, sh);
  const double alpha = rz / pq;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    scal[kPQ] = pq;
    scal[kRZ] = rz;
    scal[kAlpha] = alpha;
  }
  double rzn = 0.0, rr = 0.0;
  const XcdSpan sp = xcd_span(nnodes);
  for (int n = sp.lo + sp.j * kT + threadIdx.x; n < sp.hi; n += sp.bx * kT) {
    const size_t d = 3 * (size_t)n;
    double rv[3], zv[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      x[d + i] += alpha * p[d + i];
      rv[i] = r[d + i] - alpha * q[d + i];
      r[d + i] = rv[i];
    }
    dinv3(dinv + 9 * (size_t)n, rv, zv);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      z[d + i] = zv[i];
      rzn += rv[i] * zv[i];
      rr += rv[i] * rv[i];
    }
  }
  const double t0 = block_reduce_all<kT>(rzn, sh);
  const double t1 = block_reduce_all<kT>(rr, sh);
  if (threadIdx.x == 0) {
    P_rz_next[blockIdx.x] = t0;
    P_rr[blockIdx.x] = t1;
  }
}

// Start of PCG: z = Dinv r, p = z; partials of r.z, r.r and b.b.
__global__ __launch_bounds__(kT) void k_pcg_init(const double* __restrict__ r, double* __restrict__ z,
                                                 double* __restrict__ p, const double* __restrict__ b,
                                                 const double* __restrict__ dinv, int nnodes,
                                                 double* __restrict__ P_rz, double* __restrict__ P_rr,
                                                 double* __restrict__ P_bb) {
  __shared__ double sh[kT / 64];
  double rz = 0.0, rr = 0.0, bb = 0.0;
  const XcdSpan sp = xcd_span(nnodes);
  for (int n = sp.lo + sp.j * kT + threadIdx.x; n < sp.hi; n += sp.bx * kT) {
    const size_t d = 3 * (size_t)n;
    double rv[3] = {r[d], r[d + 1], r[d + 2]}, zv[3];
    dinv3(dinv + 9 * (size_t)n, rv, zv);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      z[d + i] = zv[i];
      p[d + i] = zv[i];
      rz += rv[i] * zv[i];
      rr += rv[i] * rv[i];
      bb += b[d + i] * b[d + i];
    }
  }
  const double t0 = block_reduce_all<kT>(rz, sh);
  const double t1 = block_reduce_all<kT>(rr, sh);
  const double t2 = block_reduce_all<kT>(bb, sh);
  if (threadIdx.x == 0) {
    P_rz[blockIdx.x] = t0;
    P_rr[blockIdx.x] = t1;
    P_bb[blockIdx.x] = t2;
  }
}

// per-group partials of a.b
__global__ __launch_bounds__(kT) void k_dot(const double* __restrict__ a, const double* __restrict__ b, size_t n,
                                            double* __restrict__ P) {
  __shared__ double sh[kT / 64];
  double v = 0.0;
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) v += a[i] * b[i];
  const double t = block_reduce_all<kT>(v, sh);
  if (threadIdx.x == 0) P[blockIdx.x] = t;
}

// fixed-order sum of n partials into scal[slot] (one work-group)
__global__ __launch_bounds__(kT) void k_sum(const double* partials, int n, double* scal, int slot) {
  __shared__ double sh[kT / 64];
  const double t = sum_partials_all<kT>(partials, n, sh);
  if (threadIdx.x == 0) scal[slot] = t;
}


// 3x3 block-diagonal inverse (cofactors)
__global__ __launch_bounds__(kT) void k_diag_inv(const double* vals, const int* diag_slot, double* dinv,
                                                 int nnodes) {
  for (int n = blockIdx.x * kT + threadIdx.x; n < nnodes; n += gridDim.x * kT) {
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
}

// ------------------------------------------------------------------ loads and BCs
// y[3n+i] += s * f[3k+i] for the loaded nodes (sparse traction vector).
__global__ __launch_bounds__(kT) void k_add_load(double* y, const int* nodes, const double* f, int n, double s) {
  for (int k = blockIdx.x * kT + threadIdx.x; k < n; k += gridDim.x * kT) {
    const size_t d = 3 * (size_t)nodes[k];
    y[d] += s * f[3 * k];
    y[d + 1] += s * f[3 * k + 1];
    y[d + 2] += s * f[3 * k + 2];
  }
}

__global__ __launch_bounds__(kT) void k_scatter_add(double* y, const int64_t* idx, const double* v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) y[idx[i]] += v[i];
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
  for (int k = blockIdx.x * kT + threadIdx.x; k < nbc; k += gridDim.x * kT) {
    const double v = scale * bcv[k];
    if (v == 0.0) continue;
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
}

// applyBlockConstraints first loop, matrix part: BC row -> unit diagonal, zero
// elsewhere; the symmetric (col,row) entries zeroed too.
__global__ __launch_bounds__(kT) void k_bc_zero(double* vals, const int* rowmap, const int* col,
                                                const int* dofs, int nbc) {
  for (int k = blockIdx.x * kT + threadIdx.x; k < nbc; k += gridDim.x * kT) {
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
}

// second loop: rhs(dof) = scale * value
__global__ __launch_bounds__(kT) void k_bc_set(double* rhs, const int* dofs, const double* bcv, int nbc,
                                               double scale) {
  for (int k = blockIdx.x * kT + threadIdx.x; k < nbc; k += gridDim.x * kT) rhs[dofs[k]] = scale * bcv[k];
}

// zero the BC dofs of a vector (homogeneous constraint on a correction)
__global__ __launch_bounds__(kT) void k_bc_zero_vec(double* y, const int* dofs, int nbc) {
  for (int k = blockIdx.x * kT + threadIdx.x; k < nbc; k += gridDim.x * kT) y[dofs[k]] = 0.0;
}

}  // namespace pa
