// Geometric multigrid kernels (see pa_mg.hpp).  Box node grids: node (i, j, k) of a
// level has grid index (i * ny + j) * nz + k; coarse node (I, J, K) sits on fine node
// (2I, 2J, 2K).  Trilinear prolongation: a fine coordinate 2I contributes weight 1 to
// coarse I, 2I +- 1 weight 1/2 to each neighbour; restriction is its transpose.
#include "pa_mg.hpp"
#include "pa_kernels.hpp"

#include <algorithm>

#include "pa_reduce.hpp"
#include "pa_solver.hpp"
#include "pa_spmv.hpp"

namespace pa {

namespace {

constexpr int kT = 256;

int blocks_of(size_t n, int cap = 4096) {
  size_t g = (n + kT - 1) / kT;
  return (int)(g < 1 ? 1 : (g > (size_t)cap ? cap : g));
}

__device__ __forceinline__ void dinv3(const double* __restrict__ D, const double r[3], double z[3]) {
  z[0] = D[0] * r[0] + D[1] * r[1] + D[2] * r[2];
  z[1] = D[3] * r[0] + D[4] * r[1] + D[5] * r[2];
  z[2] = D[6] * r[0] + D[7] * r[1] + D[8] * r[2];
}

__global__ __launch_bounds__(kT) void k_mg_restrict(const double* __restrict__ rf, double* __restrict__ rc,
                                                    const int32_t* __restrict__ cgrid,
                                                    const int32_t* __restrict__ fnode, BoxDims fd,
                                                    const uint8_t* __restrict__ cbc, int ncn,
                                                    const double* __restrict__ cdinv, double inv_theta,
                                                    double* __restrict__ cx, double* __restrict__ cd) {
  for (int c = blockIdx.x * kT + threadIdx.x; c < ncn; c += gridDim.x * kT) {
    const int I = cgrid[3 * c], J = cgrid[3 * c + 1], K = cgrid[3 * c + 2];
    // branch-free over the 27 fine points: one outside the fine grid reads the coarse
    // node's own fine point with weight 0 (adds an exact zero), so all lookups issue at once
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int a = -1; a <= 1; ++a)
#pragma unroll
      for (int b = -1; b <= 1; ++b)
#pragma unroll
        for (int e = -1; e <= 1; ++e) {
          const int i = 2 * I + a, j = 2 * J + b, k = 2 * K + e;
          const bool in = (unsigned)i < (unsigned)fd.nx && (unsigned)j < (unsigned)fd.ny && (unsigned)k < (unsigned)fd.nz;
          const double w = in ? (a ? 0.5 : 1.0) * (b ? 0.5 : 1.0) * (e ? 0.5 : 1.0) : 0.0;
          const size_t f = 3 * (size_t)fnode[((size_t)(in ? i : 2 * I) * fd.ny + (in ? j : 2 * J)) * fd.nz +
                                             (in ? k : 2 * K)];
          s0 += w * rf[f];
          s1 += w * rf[f + 1];
          s2 += w * rf[f + 2];
        }
    const uint32_t bm = cbc[c];
    const double v[3] = {(bm & 1u) ? 0.0 : s0, (bm & 2u) ? 0.0 : s1, (bm & 4u) ? 0.0 : s2};
    const size_t o = 3 * (size_t)c;
    rc[o] = v[0];
    rc[o + 1] = v[1];
    rc[o + 2] = v[2];
    if (cx) {
      double z[3];
      dinv3(cdinv + 9 * (size_t)c, v, z);
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        cx[o + i] = inv_theta * z[i];
        cd[o + i] = inv_theta * z[i];
      }
    }
  }
}

__global__ __launch_bounds__(kT) void k_mg_prolong_add(const double* __restrict__ ec, double* __restrict__ xf,
                                                       const int32_t* __restrict__ fgrid,
                                                       const int32_t* __restrict__ cnode, BoxDims cd,
                                                       const uint8_t* __restrict__ fbc, int nfn) {
  for (int n = blockIdx.x * kT + threadIdx.x; n < nfn; n += gridDim.x * kT) {
    const int i = fgrid[3 * n], j = fgrid[3 * n + 1], k = fgrid[3 * n + 2];
    const int I0 = i >> 1, J0 = j >> 1, K0 = k >> 1;
    const int nI = (i & 1) ? 2 : 1, nJ = (j & 1) ? 2 : 1, nK = (k & 1) ? 2 : 1;
    const double wI = (i & 1) ? 0.5 : 1.0, wJ = (j & 1) ? 0.5 : 1.0, wK = (k & 1) ? 0.5 : 1.0;
    // branch-free over the 8 candidates: an absent one reads coarse point (I0, J0, K0) with
    // weight 0 (adds an exact zero)
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const bool in = a < nI && b < nJ && e < nK;
          const size_t c = 3 * (size_t)cnode[((size_t)(I0 + (in ? a : 0)) * cd.ny + (J0 + (in ? b : 0))) * cd.nz +
                                             (K0 + (in ? e : 0))];
          const double w = in ? wI * wJ * wK : 0.0;
          s0 += w * ec[c];
          s1 += w * ec[c + 1];
          s2 += w * ec[c + 2];
        }
    const uint32_t bm = fbc[n];
    const size_t d = 3 * (size_t)n;
    if (!(bm & 1u)) xf[d] += s0;
    if (!(bm & 2u)) xf[d + 1] += s1;
    if (!(bm & 4u)) xf[d + 2] += s2;
  }
}

__global__ __launch_bounds__(kT) void k_mg_wrestrict(const double* __restrict__ wf, double* __restrict__ wc,
                                                     const int32_t* __restrict__ cgrid,
                                                     const int32_t* __restrict__ fnode, BoxDims fd, int ncn) {
  for (int c = blockIdx.x * kT + threadIdx.x; c < ncn; c += gridDim.x * kT) {
    const int I = cgrid[3 * c], J = cgrid[3 * c + 1], K = cgrid[3 * c + 2];
    double s = 0.0, ws = 0.0;
    for (int a = -1; a <= 1; ++a) {
      const int i = 2 * I + a;
      if (i < 0 || i >= fd.nx) continue;
      for (int b = -1; b <= 1; ++b) {
        const int j = 2 * J + b;
        if (j < 0 || j >= fd.ny) continue;
        for (int e = -1; e <= 1; ++e) {
          const int k = 2 * K + e;
          if (k < 0 || k >= fd.nz) continue;
          const double w = (a ? 0.5 : 1.0) * (b ? 0.5 : 1.0) * (e ? 0.5 : 1.0);
          s += w * wf[fnode[((size_t)i * fd.ny + j) * fd.nz + k]];
          ws += w;
        }
      }
    }
    wc[c] = s / ws;
  }
}

__global__ __launch_bounds__(kT) void k_simp_w(const double* z, double* w, int n, Material m) {
  for (int i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) w[i] = penalty_w(m, z[i]);
}

__global__ __launch_bounds__(kT) void k_cheb_first(double* __restrict__ x, double* __restrict__ d,
                                                   const double* __restrict__ r, const double* __restrict__ dinv,
                                                   double inv_theta, int nn, const double* __restrict__ t) {
  for (int n = blockIdx.x * kT + threadIdx.x; n < nn; n += gridDim.x * kT) {
    const size_t o = 3 * (size_t)n;
    double rv[3] = {r[o], r[o + 1], r[o + 2]};
    if (t) {
#pragma unroll
      for (int i = 0; i < 3; ++i) rv[i] -= t[o + i];
    }
    double zv[3];
    dinv3(dinv + 9 * (size_t)n, rv, zv);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double v = inv_theta * zv[i];
      if (d) d[o + i] = v;
      x[o + i] = v;
    }
  }
}

__global__ __launch_bounds__(kT) void k_cheb_next(double* __restrict__ x, double* __restrict__ d,
                                                  const double* __restrict__ r, const double* __restrict__ Kx,
                                                  const double* __restrict__ dinv, double c1, double c2, int nn) {
  for (int n = blockIdx.x * kT + threadIdx.x; n < nn; n += gridDim.x * kT) {
    const size_t o = 3 * (size_t)n;
    const double rv[3] = {r[o] - Kx[o], r[o + 1] - Kx[o + 1], r[o + 2] - Kx[o + 2]};
    double zv[3];
    dinv3(dinv + 9 * (size_t)n, rv, zv);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double v = c1 * d[o + i] + c2 * zv[i];
      d[o + i] = v;
      x[o + i] += v;
    }
  }
}

__global__ __launch_bounds__(kT) void k_sub(double* out, const double* a, const double* b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) out[i] = a[i] - b[i];
}

__global__ __launch_bounds__(kT) void k_add(double* y, const double* x, size_t n) {
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) y[i] += x[i];
}

__global__ __launch_bounds__(kT) void k_add2(double* out, const double* a, const double* b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) out[i] = a[i] + b[i];
}

__global__ __launch_bounds__(kT) void k_dinv_apply(double* __restrict__ y, const double* __restrict__ dinv,
                                                   const double* __restrict__ x, int nn) {
  for (int n = blockIdx.x * kT + threadIdx.x; n < nn; n += gridDim.x * kT) {
    const size_t o = 3 * (size_t)n;
    const double xv[3] = {x[o], x[o + 1], x[o + 2]};
    double zv[3];
    dinv3(dinv + 9 * (size_t)n, xv, zv);
    y[o] = zv[0];
    y[o + 1] = zv[1];
    y[o + 2] = zv[2];
  }
}

// one wavefront per row, fixed-order lane sums
__global__ __launch_bounds__(kT) void k_dense_gemv(const double* __restrict__ A, const double* __restrict__ x,
                                                   double* __restrict__ y, int n) {
  const int lane = threadIdx.x & 63;
  for (int row = (blockIdx.x * kT + threadIdx.x) / 64; row < n; row += gridDim.x * (kT / 64)) {
    double s = 0.0;
    for (int c = lane; c < n; c += 64) s += A[(size_t)row * n + c] * x[c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) y[row] = s;
  }
}

__global__ __launch_bounds__(kT) void k_mg_update(double* __restrict__ x, double* __restrict__ r,
                                                  const double* __restrict__ p, const double* __restrict__ q,
                                                  const double* __restrict__ scal, int nn, double* __restrict__ P_rr) {
  __shared__ double sh[kT / 64];
  const double alpha = scal[kAlpha];
  double rr = 0.0;
  const XcdSpan sp = xcd_span(nn);
  for (int n = sp.lo + sp.j * kT + threadIdx.x; n < sp.hi; n += sp.bx * kT) {
    const size_t d = 3 * (size_t)n;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      x[d + i] += alpha * p[d + i];
      const double rv = r[d + i] - alpha * q[d + i];
      r[d + i] = rv;
      rr += rv * rv;
    }
  }
  const double t = block_reduce_all<kT>(rr, sh);
  if (threadIdx.x == 0) P_rr[blockIdx.x] = t;
}

__global__ __launch_bounds__(kT) void k_mg_pupdate(double* __restrict__ p, const double* __restrict__ z,
                                                   const double* __restrict__ scal, size_t n) {
  const double beta = scal[kBeta];
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT)
    p[i] = fma(beta, p[i], z[i]);
}

template <int MODE>
__device__ __forceinline__ void mg_row_epilogue(const MgRowArgs& A, int row, const double acc[3]) {
  {
    const double* __restrict__ x = A.x;
    const size_t o = 3 * (size_t)row;
    const double t[3] = {A.r[o] - acc[0], A.r[o + 1] - acc[1], A.r[o + 2] - acc[2]};
    if constexpr (MODE == kMgRes) {
#pragma unroll
      for (int i = 0; i < 3; ++i) A.out[o + i] = t[i];
    } else if constexpr (MODE == kMgResCheb0) {
      double z[3];
      dinv3(A.dinv + 9 * (size_t)row, t, z);
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        A.out[o + i] = t[i];
        A.y[o + i] = A.inv_theta * z[i];
        A.d[o + i] = A.inv_theta * z[i];
      }
    } else {
      double z[3];
      dinv3(A.dinv + 9 * (size_t)row, t, z);
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const double dn = A.c1 * A.d[o + i] + A.c2 * z[i];
        A.d[o + i] = dn;
        if constexpr (MODE == kMgCheb) A.out[o + i] = x[o + i] + dn;
        else A.out[o + i] = A.base[o + i] + x[o + i] + dn;
      }
    }
  }
}

template <int MODE, typename V>
__global__ __launch_bounds__(kT) void k_mg_rows(MgRowArgs A) {
  constexpr int kRows = kT / kLanesPerRow;
  const int lane = threadIdx.x % kLanesPerRow;
  const XcdSpan sp = xcd_span(A.nrows);
  const double* __restrict__ x = A.x;
  const V* __restrict__ vals;
  if constexpr (sizeof(V) == 4) vals = A.vals32;
  else vals = A.vals;
  auto X = [&](int c, int j) { return x[3 * (size_t)c + j]; };
  for (int row = sp.lo + sp.j * kRows + threadIdx.x / kLanesPerRow; row < sp.hi; row += sp.bx * kRows) {
    double acc[3];
    spmv_row(vals, A.rowmap, A.col, row, lane, X, acc);
    if (lane == 0) mg_row_epilogue<MODE>(A, row, acc);
  }
}

__global__ __launch_bounds__(kT) void k_to_f32(const double* __restrict__ a, float* __restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) b[i] = (float)a[i];
}

__global__ __launch_bounds__(kT) void k_scale_rsqrt(double* __restrict__ y, size_t n, const double* __restrict__ scal,
                                                    int slot) {
  const double f = 1.0 / sqrt(scal[slot]);
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) y[i] *= f;
}

}  // namespace

hipError_t launch_mg_restrict(const double* rf, double* rc, const int32_t* cgrid, const int32_t* fnode_of_grid,
                              BoxDims fd, const uint8_t* cbcmask, int ncn, hipStream_t s, const double* cdinv,
                              double inv_theta, double* cx, double* cd) {
  k_mg_restrict<<<blocks_of(ncn), kT, 0, s>>>(rf, rc, cgrid, fnode_of_grid, fd, cbcmask, ncn, cdinv, inv_theta, cx,
                                              cd);
  return hipGetLastError();
}
hipError_t launch_mg_rows(int mode, const MgRowArgs& a, hipStream_t s) {
  const int grid = red_grid((size_t)a.nrows, kT / kLanesPerRow);
  if (a.vals32) {
    switch (mode) {
      case kMgRes: k_mg_rows<kMgRes, float><<<grid, kT, 0, s>>>(a); break;
      case kMgResCheb0: k_mg_rows<kMgResCheb0, float><<<grid, kT, 0, s>>>(a); break;
      case kMgCheb: k_mg_rows<kMgCheb, float><<<grid, kT, 0, s>>>(a); break;
      case kMgChebAdd: k_mg_rows<kMgChebAdd, float><<<grid, kT, 0, s>>>(a); break;
      default: return hipErrorInvalidValue;
    }
  } else {
    switch (mode) {
      case kMgRes: k_mg_rows<kMgRes, double><<<grid, kT, 0, s>>>(a); break;
      case kMgResCheb0: k_mg_rows<kMgResCheb0, double><<<grid, kT, 0, s>>>(a); break;
      case kMgCheb: k_mg_rows<kMgCheb, double><<<grid, kT, 0, s>>>(a); break;
      case kMgChebAdd: k_mg_rows<kMgChebAdd, double><<<grid, kT, 0, s>>>(a); break;
      default: return hipErrorInvalidValue;
    }
  }
  return hipGetLastError();
}
hipError_t launch_to_f32(const double* a, float* b, size_t n, hipStream_t s) {
  k_to_f32<<<blocks_of(n), kT, 0, s>>>(a, b, n);
  return hipGetLastError();
}
hipError_t launch_scale_rsqrt(double* y, size_t n, const double* scal, int slot, hipStream_t s) {
  k_scale_rsqrt<<<blocks_of(n), kT, 0, s>>>(y, n, scal, slot);
  return hipGetLastError();
}
hipError_t launch_mg_prolong_add(const double* ec, double* xf, const int32_t* fgrid, const int32_t* cnode_of_grid,
                                 BoxDims cd, const uint8_t* fbcmask, int nfn, hipStream_t s) {
  k_mg_prolong_add<<<blocks_of(nfn), kT, 0, s>>>(ec, xf, fgrid, cnode_of_grid, cd, fbcmask, nfn);
  return hipGetLastError();
}
hipError_t launch_mg_wrestrict(const double* wf, double* wc, const int32_t* cgrid, const int32_t* fnode_of_grid,
                               BoxDims fd, int ncn, hipStream_t s) {
  k_mg_wrestrict<<<blocks_of(ncn), kT, 0, s>>>(wf, wc, cgrid, fnode_of_grid, fd, ncn);
  return hipGetLastError();
}
hipError_t launch_simp_w(const double* z, double* w, int n, const Material& m, hipStream_t s) {
  k_simp_w<<<blocks_of(n), kT, 0, s>>>(z, w, n, m);
  return hipGetLastError();
}
hipError_t launch_cheb_first(double* x, double* d, const double* r, const double* dinv, double inv_theta, int nn,
                             hipStream_t s, const double* t) {
  k_cheb_first<<<blocks_of(nn), kT, 0, s>>>(x, d, r, dinv, inv_theta, nn, t);
  return hipGetLastError();
}
hipError_t launch_cheb_next(double* x, double* d, const double* r, const double* Kx, const double* dinv, double c1,
                            double c2, int nn, hipStream_t s) {
  k_cheb_next<<<blocks_of(nn), kT, 0, s>>>(x, d, r, Kx, dinv, c1, c2, nn);
  return hipGetLastError();
}
hipError_t launch_sub(double* out, const double* a, const double* b, size_t n, hipStream_t s) {
  k_sub<<<blocks_of(n), kT, 0, s>>>(out, a, b, n);
  return hipGetLastError();
}
hipError_t launch_add(double* y, const double* x, size_t n, hipStream_t s) {
  k_add<<<blocks_of(n), kT, 0, s>>>(y, x, n);
  return hipGetLastError();
}
hipError_t launch_add2(double* out, const double* a, const double* b, size_t n, hipStream_t s) {
  k_add2<<<blocks_of(n), kT, 0, s>>>(out, a, b, n);
  return hipGetLastError();
}
hipError_t launch_dinv_apply(double* y, const double* dinv, const double* x, int nn, hipStream_t s) {
  k_dinv_apply<<<blocks_of(nn), kT, 0, s>>>(y, dinv, x, nn);
  return hipGetLastError();
}
hipError_t launch_dense_gemv(const double* A, const double* x, double* y, int n, hipStream_t s) {
  k_dense_gemv<<<std::max(1, std::min(1024, (n + 3) / 4)), kT, 0, s>>>(A, x, y, n);
  return hipGetLastError();
}
hipError_t launch_mg_update(double* x, double* r, const double* p, const double* q, const double* scal, int nn,
                            double* P_rr, int grid, hipStream_t s) {
  k_mg_update<<<grid, kT, 0, s>>>(x, r, p, q, scal, nn, P_rr);
  return hipGetLastError();
}
hipError_t launch_mg_pupdate(double* p, const double* z, const double* scal, size_t n, hipStream_t s) {
  k_mg_pupdate<<<blocks_of(n), kT, 0, s>>>(p, z, scal, n);
  return hipGetLastError();
}

}  // namespace pa
