// Mixed-precision solve: FP32 matrix-free operator inside an FP64 Jacobi-PCG (FP64
// vectors and reductions, residual replaced by the true FP64 residual f - K u).
#include "pa_mf.hpp"
#include "pa_fixed.hpp"
#include "pa_reduce.hpp"
#include "pa_solver.hpp"

namespace pa {

namespace {

constexpr int kT = 256;

__device__ __forceinline__ int xcd_map(int b, int n) {
  const int q = n / kNumXcd, r = n % kNumXcd;
  const int x = b % kNumXcd, k = b / kNumXcd;
  return x * q + (x < r ? x : r) + k;
}

__device__ __forceinline__ float powi_f(float x, float p, int pint) {
  if (pint > 0) {
    float r = x;
    for (int k = 1; k < pint; ++k) r *= x;
    return r;
  }
  return powf(x, p);
}

template <typename T>
__device__ __forceinline__ void gradients_t(const T x[4][3], T g[4][3], T& vol) {
  const T j11 = x[0][0] - x[3][0], j12 = x[1][0] - x[3][0], j13 = x[2][0] - x[3][0];
  const T j21 = x[0][1] - x[3][1], j22 = x[1][1] - x[3][1], j23 = x[2][1] - x[3][1];
  const T j31 = x[0][2] - x[3][2], j32 = x[1][2] - x[3][2], j33 = x[2][2] - x[3][2];
  const T det = j11 * j22 * j33 + j12 * j23 * j31 + j13 * j21 * j32 - j11 * j23 * j32 - j12 * j21 * j33 -
                j13 * j22 * j31;
  const T inv = T(1) / det;
  g[0][0] = (j22 * j33 - j23 * j32) * inv;
  g[0][1] = (j13 * j32 - j12 * j33) * inv;
  g[0][2] = (j12 * j23 - j13 * j22) * inv;
  g[1][0] = (j23 * j31 - j21 * j33) * inv;
  g[1][1] = (j11 * j33 - j13 * j31) * inv;
  g[1][2] = (j13 * j21 - j11 * j23) * inv;
  g[2][0] = (j21 * j32 - j22 * j31) * inv;
  g[2][1] = (j12 * j31 - j11 * j32) * inv;
  g[2][2] = (j11 * j22 - j12 * j21) * inv;
#pragma unroll
  for (int d = 0; d < 3; ++d) g[3][d] = -(g[0][d] + g[1][d] + g[2][d]);
  vol = fabs(det) * (T(1) / T(6));
}

// q = K p with the PCG direction update fused in; one work-group per chunk.  The
// vectors are FP64; the element operator is FP32: the node table (FP32 coordinates +
// density) and p = z + beta p_prev at the table's nodes (rounded to FP32, constrained
// dofs 0) are staged in LDS, the element loop computes K_e p_e in FP32 and adds it into
// 64-bit fixed-point row accumulators (scale from the chunk's largest |p|: the sums do
// not depend on the order of the adds), and the rows are written back with the
// chunk's p.q partial.
// The chunk's own rows of p_cur are written in FP64 (unrounded).
template <int NT>
__global__ __launch_bounds__(NT) void k_mf_apply(MfArgs A) {
  // LDS: own-row p and row accumulators (FP64, 3 x even(max_rows) each: multiples of
  // 16 bytes), node table, p at the table's nodes (FP32).  FP64 LDS atomics run at
  // several times the rate of FP32 ones on this part, and add no rounding of their own.
  extern __shared__ double smem_d[];
  const int mr2 = (A.max_rows + 1) & ~1;
  double* pown = smem_d;
  double* acc = pown + 3 * mr2;
  float4* nd4 = reinterpret_cast<float4*>(acc + 3 * mr2);
  float* xl = reinterpret_cast<float*>(nd4 + A.max_nodes);
  __shared__ double sh[NT / 64];
  __shared__ double red[NT / 64];

  const double beta = A.first ? 0.0 : A.scal[kBeta];
  const int c = xcd_map(blockIdx.x, A.nchunks);
  const int r0 = A.chunk_rows[c], r1 = A.chunk_rows[c + 1], nr = r1 - r0;
  const int n0 = A.node_offs[c], nnode = A.node_offs[c + 1] - n0;
  double pmax = 0.0;
  for (int k = threadIdx.x; k < nnode; k += NT) {
    const int gid = A.nodes[n0 + k];
    nd4[k] = A.Xz[gid];
    const uint32_t bm = A.bcmask[gid];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const size_t d = 3 * (size_t)gid + i;
      double v = A.first ? A.z[d] : fma(beta, A.p_prev[d], A.z[d]);
      v = (bm >> i) & 1u ? 0.0 : v;
      xl[3 * k + i] = (float)v;
      pmax = max_nan(pmax, fabs((double)(float)v));
      if (k < nr) {  // rows come first in the table
        A.p_cur[d] = v;
        pown[3 * k + i] = v;
      }
    }
  }
  for (int k = threadIdx.x; k < 3 * nr; k += NT) acc[k] = 0.0;
  block_max_write<NT>(pmax, red);
  __syncthreads();
  // |contribution| <= vol |grad N|^2 * 3 * (|c00| + 2|c01| + |c33|) * 8 max|p| for weights
  // <= 1, plus FP32 rounding (the factor 2)
  const double S = fixed_scale_54(48.0 * (fabs(A.m.c00) + 2.0 * fabs(A.m.c01) + fabs(A.m.c33)) *
                                  (double)A.gbound[c] * block_max_read<NT>(red));

  const float c00 = (float)A.m.c00, c01 = (float)A.m.c01, c33 = (float)A.m.c33;
  const float rmin = (float)A.m.rmin, pp = (float)A.m.p;
  const int l0 = A.list_offs[c], l1 = A.list_offs[c + 1];
  for (int k = l0 + threadIdx.x; k < l1; k += NT) {
    const uint2 lw = reinterpret_cast<const uint2*>(A.ent_ln)[k];
    const int ln[4] = {(int)(lw.x & 0xFFFFu), (int)(lw.x >> 16), (int)(lw.y & 0xFFFFu), (int)(lw.y >> 16)};
    float x[4][3], g[4][3], u[4][3], vol, rho = 0.0f;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const float4 p = nd4[ln[a]];
      x[a][0] = p.x;
      x[a][1] = p.y;
      x[a][2] = p.z;
      rho += p.w;
#pragma unroll
      for (int i = 0; i < 3; ++i) u[a][i] = xl[3 * ln[a] + i];
    }
    gradients_t(x, g, vol);
    rho *= 0.25f;
    const float vw = vol * (A.m.model == kPenaltySimp ? rmin + (1.0f - rmin) * powi_f(rho, pp, A.m.pint)
                                                      : (float)penalty_w(A.m, (double)rho));
    float e[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      e[0] += u[a][0] * g[a][0];
      e[1] += u[a][1] * g[a][1];
      e[2] += u[a][2] * g[a][2];
      e[3] += u[a][2] * g[a][1] + u[a][1] * g[a][2];
      e[4] += u[a][2] * g[a][0] + u[a][0] * g[a][2];
      e[5] += u[a][1] * g[a][0] + u[a][0] * g[a][1];
    }
    const float s0 = vw * (c00 * e[0] + c01 * e[1] + c01 * e[2]);
    const float s1 = vw * (c01 * e[0] + c00 * e[1] + c01 * e[2]);
    const float s2 = vw * (c01 * e[0] + c01 * e[1] + c00 * e[2]);
    const float s3 = vw * c33 * e[3], s4 = vw * c33 * e[4], s5 = vw * c33 * e[5];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      if (ln[a] < nr) {
        double* dst = &acc[3 * ln[a]];
        lds_add_fixed64(dst + 0, (double)(s0 * g[a][0] + s5 * g[a][1] + s4 * g[a][2]), S);
        lds_add_fixed64(dst + 1, (double)(s5 * g[a][0] + s1 * g[a][1] + s3 * g[a][2]), S);
        lds_add_fixed64(dst + 2, (double)(s4 * g[a][0] + s3 * g[a][1] + s2 * g[a][2]), S);
      }
    }
  }
  __syncthreads();
  double dot = 0.0;
  const double inv = 1.0 / S;  // a power of two: exact (NaN for a non-finite p)
  for (int k = threadIdx.x; k < 3 * nr; k += NT) {
    const size_t gd = 3 * (size_t)r0 + k;
    const double pv = pown[k];
    const double av = S == S ? fixed_value(acc[k], inv) : NAN;
    const double yv = (A.bcmask[r0 + k / 3] >> (k % 3)) & 1u ? pv : av;
    A.y[gd] = yv;
    dot += pv * yv;
  }
  const double t = block_reduce_all<NT>(dot, sh);
  if (threadIdx.x == 0) A.partials[c] = t;
}

// Diagonal blocks K_rr (FP64) accumulated per owned vertex, then inverted per row.
template <int NT>
__global__ __launch_bounds__(NT) void k_mf_diag(DiagArgs A) {
  extern __shared__ double dsm[];
  double* D = dsm;  // 9 per row
  const int c = xcd_map(blockIdx.x, A.nchunks);
  const int r0 = A.chunk_rows[c], r1 = A.chunk_rows[c + 1], nr = r1 - r0;
  const int n0 = A.node_offs[c];
  for (int k = threadIdx.x; k < 9 * nr; k += NT) D[k] = 0.0;
  __syncthreads();
  // fixed-point sums (order-independent): |contribution| <= vol |grad N|^2 (|c01| + 2|c33|)
  const double S = fixed_scale_54((double)A.gbound[c] * (fabs(A.m.c01) + 2.0 * fabs(A.m.c33)));
  const int l0 = A.list_offs[c], l1 = A.list_offs[c + 1];
  for (int k = l0 + threadIdx.x; k < l1; k += NT) {
    const uint2 lw = reinterpret_cast<const uint2*>(A.ent_ln)[k];
    const int ln[4] = {(int)(lw.x & 0xFFFFu), (int)(lw.x >> 16), (int)(lw.y & 0xFFFFu), (int)(lw.y >> 16)};
    double x[4][3], g[4][3], vol, rho = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const size_t gid = (size_t)A.nodes[n0 + ln[a]];
#pragma unroll
      for (int d = 0; d < 3; ++d) x[a][d] = A.X[3 * gid + d];
      rho += A.z[gid];
    }
    gradients_t(x, g, vol);
    rho = rho / 4.0;
    const double vw = vol * penalty_w(A.m, rho);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      if (ln[a] >= nr) continue;
      const double gg = g[a][0] * g[a][0] + g[a][1] * g[a][1] + g[a][2] * g[a][2];
      double* dst = &D[9 * ln[a]];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          double v = (A.m.c01 + A.m.c33) * g[a][i] * g[a][j];
          if (i == j) v += A.m.c33 * gg;
          lds_add_fixed64(dst + 3 * i + j, vw * v, S);
        }
    }
  }
  __syncthreads();
  const double inv = 1.0 / S;
  for (int lr = threadIdx.x; lr < nr; lr += NT) {
    double a[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) a[q] = fixed_value(D[9 * lr + q], inv);
    block_inv3_masked(a, A.bcmask[r0 + lr], A.dinv + 9 * (size_t)(r0 + lr));
  }
}

__device__ __forceinline__ void dinv3(const double* __restrict__ D, const double r[3], double z[3]) {
  z[0] = D[0] * r[0] + D[1] * r[1] + D[2] * r[2];
  z[1] = D[3] * r[0] + D[4] * r[1] + D[5] * r[2];
  z[2] = D[6] * r[0] + D[7] * r[1] + D[8] * r[2];
}

// z = Dinv r; partials of r.z
__global__ __launch_bounds__(kT) void k_pcgm_init(const double* __restrict__ r, double* __restrict__ z,
                                                  const double* __restrict__ dinv, int nn,
                                                  double* __restrict__ P_rz) {
  __shared__ double sh[kT / 64];
  double rz = 0.0;
  const XcdSpan sp = xcd_span(nn);
  for (int n = sp.lo + sp.j * kT + threadIdx.x; n < sp.hi; n += sp.bx * kT) {
    const size_t d = 3 * (size_t)n;
    const double rv[3] = {r[d], r[d + 1], r[d + 2]};
    double zv[3];
    dinv3(dinv + 9 * (size_t)n, rv, zv);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      z[d + i] = zv[i];
      rz += rv[i] * zv[i];
    }
  }
  const double t = block_reduce_all<kT>(rz, sh);
  if (threadIdx.x == 0) P_rz[blockIdx.x] = t;
}

constexpr int kFin = 1024;  // threads of the one-work-group finishers

__global__ __launch_bounds__(kFin) void k_pcgm_alpha(const double* partials, int n, double* scal) {
  __shared__ double sh[kFin / 64];
  const double pq = sum_partials_wide<kFin>(partials, n, sh);
  if (threadIdx.x == 0) {
    scal[kPQ] = pq;
    scal[kAlpha] = scal[kRZ] / pq;
  }
}

// x += alpha p; r -= alpha q; z = Dinv r (alpha = scal[kAlpha]); partials of r.z, r.r
__global__ __launch_bounds__(kT) void k_pcgm_update(double* __restrict__ x, double* __restrict__ r,
                                                    double* __restrict__ z, const double* __restrict__ p,
                                                    const double* __restrict__ q, const double* __restrict__ dinv,
                                                    int nn, const double* __restrict__ scal,
                                                    double* __restrict__ P_rz, double* __restrict__ P_rr) {
  __shared__ double sh[kT / 64];
  const double alpha = scal[kAlpha];
  double rzn = 0.0, rr = 0.0;
  const XcdSpan sp = xcd_span(nn);
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
    P_rz[blockIdx.x] = t0;
    P_rr[blockIdx.x] = t1;
  }
}

__global__ __launch_bounds__(kFin) void k_pcgm_beta(const double* P_rz, const double* P_rr, int n, double* scal,
                                                    int mode) {
  __shared__ double sh[kFin / 64];
  const double rz = sum_partials_wide<kFin>(P_rz, n, sh);
  const double rr = P_rr ? sum_partials_wide<kFin>(P_rr, n, sh) : 0.0;
  if (threadIdx.x == 0) {
    if (mode == kBetaStep) {
      scal[kBeta] = rz / scal[kRZ];
      scal[kRZPrev] = scal[kRZ];
    } else if (mode == kBetaReplace) {
      scal[kBeta] = rz / scal[kRZPrev];
    }
    scal[kRZ] = rz;
    if (P_rr) scal[kRR] = rr;
  }
}

// r = mask(b - R) (b optional); partials of r.r
__global__ __launch_bounds__(kT) void k_neg_mask(const double* __restrict__ R, const double* __restrict__ b,
                                                 const uint8_t* __restrict__ bcmask, double* __restrict__ r, int nn,
                                                 double* __restrict__ P) {
  __shared__ double sh[kT / 64];
  double rr = 0.0;
  const XcdSpan sp = xcd_span(nn);
  for (int n = sp.lo + sp.j * kT + threadIdx.x; n < sp.hi; n += sp.bx * kT) {
    const uint32_t bm = bcmask[n];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const size_t d = 3 * (size_t)n + i;
      const double v = (bm >> i) & 1u ? 0.0 : (b ? b[d] : 0.0) - R[d];
      r[d] = v;
      rr += v * v;
    }
  }
  const double t = block_reduce_all<kT>(rr, sh);
  if (threadIdx.x == 0) P[blockIdx.x] = t;
}

__global__ __launch_bounds__(kT) void k_jacobi_step(double* __restrict__ u, const double* __restrict__ Ku,
                                                    const double* __restrict__ dinv, const uint8_t* __restrict__ bcmask,
                                                    int nn, double omega) {
  for (int n = blockIdx.x * kT + threadIdx.x; n < nn; n += gridDim.x * kT) {
    const size_t d = 3 * (size_t)n;
    const double r[3] = {Ku[d], Ku[d + 1], Ku[d + 2]};
    double c[3];
    dinv3(dinv + 9 * (size_t)n, r, c);
    const uint32_t bm = bcmask[n];
#pragma unroll
    for (int i = 0; i < 3; ++i)
      if (!((bm >> i) & 1u)) u[d + i] -= omega * c[i];
  }
}

__global__ __launch_bounds__(kT) void k_make_xz(const double* X, const double* z, float4* Xz, int nn) {
  for (int n = blockIdx.x * kT + threadIdx.x; n < nn; n += gridDim.x * kT)
    Xz[n] = make_float4((float)X[3 * (size_t)n], (float)X[3 * (size_t)n + 1], (float)X[3 * (size_t)n + 2],
                        (float)z[n]);
}

int grid_of(size_t n, int cap = 1024) {
  size_t g = (n + kT - 1) / kT;
  return (int)(g < 1 ? 1 : (g > (size_t)cap ? cap : g));
}

}  // namespace

hipError_t launch_mf_apply(const MfArgs& a, hipStream_t s) {
  const size_t lds = sizeof(double) * 6 * ((a.max_rows + 1) & ~1) + sizeof(float4) * a.max_nodes +
                     sizeof(float) * 3 * a.max_nodes;
  k_mf_apply<256><<<a.nchunks, 256, lds, s>>>(a);
  return hipGetLastError();
}
hipError_t launch_mf_diag_inv(const DiagArgs& a, hipStream_t s) {
  k_mf_diag<256><<<a.nchunks, 256, sizeof(double) * 9 * a.max_rows, s>>>(a);
  return hipGetLastError();
}
hipError_t launch_pcgm_init(const double* r, double* z, const double* dinv, int nn, double* P_rz, int grid,
                            hipStream_t s) {
  k_pcgm_init<<<grid, kT, 0, s>>>(r, z, dinv, nn, P_rz);
  return hipGetLastError();
}
hipError_t launch_pcgm_alpha(const double* partials, int n, double* scal, hipStream_t s) {
  k_pcgm_alpha<<<1, kFin, 0, s>>>(partials, n, scal);
  return hipGetLastError();
}
hipError_t launch_pcgm_update(double* x, double* r, double* z, const double* p, const double* q, const double* dinv,
                              int nn, const double* scal, double* P_rz, double* P_rr, int grid, hipStream_t s) {
  k_pcgm_update<<<grid, kT, 0, s>>>(x, r, z, p, q, dinv, nn, scal, P_rz, P_rr);
  return hipGetLastError();
}
hipError_t launch_pcgm_beta(const double* P_rz, const double* P_rr, int n, double* scal, int mode, hipStream_t s) {
  k_pcgm_beta<<<1, kFin, 0, s>>>(P_rz, P_rr, n, scal, mode);
  return hipGetLastError();
}
hipError_t launch_neg_mask(const double* R, const double* b, const uint8_t* bcmask, double* r, int nn, double* P,
                           int grid, hipStream_t s) {
  k_neg_mask<<<grid, kT, 0, s>>>(R, b, bcmask, r, nn, P);
  return hipGetLastError();
}
hipError_t launch_jacobi_step(double* u, const double* Ku, const double* dinv, const uint8_t* bcmask, int nn,
                              double omega, hipStream_t s) {
  k_jacobi_step<<<grid_of(nn, 4096), kT, 0, s>>>(u, Ku, dinv, bcmask, nn, omega);
  return hipGetLastError();
}
hipError_t launch_make_xz(const double* X, const double* z, float4* Xz, int nn, hipStream_t s) {
  k_make_xz<<<grid_of(nn, 4096), kT, 0, s>>>(X, z, Xz, nn);
  return hipGetLastError();
}

}  // namespace pa
