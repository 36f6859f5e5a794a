// Deflation kernels for the deflated PCG (pa_deflate.hpp).
#include "pa_deflate.hpp"
#include "pa_reduce.hpp"

namespace pa {

namespace {
constexpr int kT = 256;
}

// Partials of V_j . x (j < k) per work-group, over an XCD-contiguous node range.
__global__ __launch_bounds__(kT) void k_multi_dot(const double* __restrict__ V, int k, size_t ld,
                                                  const double* __restrict__ x, int nnodes, double* __restrict__ P) {
  __shared__ double sh[kT / 64];
  double acc[kMaxDeflation];
#pragma unroll
  for (int j = 0; j < kMaxDeflation; ++j) acc[j] = 0.0;
  const XcdSpan sp = xcd_span(nnodes);
  for (int n = sp.lo + sp.j * kT + threadIdx.x; n < sp.hi; n += sp.bx * kT) {
    const size_t d = 3 * (size_t)n;
    const double x0 = x[d], x1 = x[d + 1], x2 = x[d + 2];
#pragma unroll
    for (int j = 0; j < kMaxDeflation; ++j)
      if (j < k) {
        const double* v = V + j * ld + d;
        acc[j] += v[0] * x0 + v[1] * x1 + v[2] * x2;
      }
  }
  for (int j = 0; j < k; ++j) {
    const double t = block_reduce_all<kT>(acc[j], sh);
    if (threadIdx.x == 0) P[(size_t)j * gridDim.x + blockIdx.x] = t;
  }
}

// mu = Einv s with s_j the fixed-order sum of the np partials of column j.
__global__ __launch_bounds__(kT) void k_defl_mu(const double* __restrict__ P, int np, int k,
                                                const double* __restrict__ Einv, double* __restrict__ mu) {
  __shared__ double sh[kT / 64];
  __shared__ double s[kMaxDeflation];
  for (int j = 0; j < k; ++j) {
    const double t = sum_partials_all<kT>(P + (size_t)j * np, np, sh);
    if (threadIdx.x == 0) s[j] = t;
  }
  __syncthreads();
  if (threadIdx.x < k) {
    double m = 0.0;
    for (int j = 0; j < k; ++j) m += Einv[threadIdx.x * k + j] * s[j];
    mu[threadIdx.x] = m;
  }
}

// y += sign * sum_j mu_j V_j
__global__ __launch_bounds__(kT) void k_multi_axpy(double* __restrict__ y, const double* __restrict__ V, int k,
                                                   size_t ld, const double* __restrict__ mu, double sign, size_t n) {
  __shared__ double m[kMaxDeflation];
  if (threadIdx.x < k) m[threadIdx.x] = sign * mu[threadIdx.x];
  __syncthreads();
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) {
    double a = y[i];
    for (int j = 0; j < k; ++j) a = fma(m[j], V[j * ld + i], a);
    y[i] = a;
  }
}

}  // namespace pa
