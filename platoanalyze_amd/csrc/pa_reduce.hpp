// Fence-free grid reductions and XCD-aware persistent grids.
//
// A reduction kernel writes one partial per work-group; the kernel that needs the
// total (the next one in the stream) sums the partials in a fixed order in every
// work-group.  Kernel boundaries make the partials visible, so no work-group ever
// issues a device-scope fence (on a multi-XCD part such a fence writes back the
// issuing XCD's L2) and the result is deterministic run to run.
#pragma once
#include <hip/hip_runtime.h>

#include "pa_kernels.hpp"

namespace pa {

// Work-groups of the persistent vector / reduction kernels (also the partial count);
// a multiple of the XCD count.
constexpr int kRedBlocks = 1024;

// Grid of a persistent kernel over `work` items handled `per_block` at a time.
inline int red_grid(size_t work, int per_block) {
  size_t g = (work + per_block - 1) / per_block;
  g = (g + kNumXcd - 1) / kNumXcd * kNumXcd;
  if (g < (size_t)kNumXcd) g = kNumXcd;
  if (g > (size_t)kRedBlocks) g = kRedBlocks;
  return (int)g;
}

// Sum of one value per thread over the work-group, returned in every thread
// (fixed order).  `sh` holds NT/64 doubles; safe to call repeatedly.
template <int NT>
__device__ __forceinline__ double block_reduce_all(double v, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int q = 0; q < NT / 64; ++q) t += sh[q];
  return t;
}

// Fixed-order total of n partials written by an earlier kernel, in every thread.
template <int NT>
__device__ __forceinline__ double sum_partials_all(const double* p, int n, double* sh) {
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += NT) v += p[i];
  return block_reduce_all<NT>(v, sh);
}

// Same, with four independent accumulators per thread (for long partial arrays read by
// a single work-group: more loads in flight).  Fixed order, deterministic.
template <int NT>
__device__ __forceinline__ double sum_partials_wide(const double* p, int n, double* sh) {
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int i = threadIdx.x;
  for (; i + 3 * NT < n; i += 4 * NT) {
    a0 += p[i];
    a1 += p[i + NT];
    a2 += p[i + 2 * NT];
    a3 += p[i + 3 * NT];
  }
  for (; i < n; i += NT) a0 += p[i];
  return block_reduce_all<NT>((a0 + a1) + (a2 + a3), sh);
}

// Contiguous share [lo, hi) of n items for the XCD that runs this work-group
// (work-group b runs on XCD b % 8), and this group's rank j among the bx groups
// of that XCD.  Items close in index (Hilbert-ordered rows, chunks) stay in one L2.
struct XcdSpan {
  int lo, hi, j, bx;
};
__device__ __forceinline__ XcdSpan xcd_span(int n) {
  XcdSpan s;
  const int x = blockIdx.x % kNumXcd;
  s.j = blockIdx.x / kNumXcd;
  s.bx = gridDim.x / kNumXcd;
  const int per = (n + kNumXcd - 1) / kNumXcd;
  s.lo = min(n, x * per);
  s.hi = min(n, s.lo + per);
  return s;
}

}  // namespace pa
