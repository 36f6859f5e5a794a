// Order-independent LDS sums: every contribution is rounded once to a multiple of 1/S
// (S a power of two chosen from a bound of the contributions) and added as a 64-bit
// integer, so the sum does not depend on the order in which lanes and waves arrive.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace pa {

// Largest power of two S with S * bound < 2^54: up to 512 contributions of magnitude
// <= bound per slot cannot overflow, and each is resolved to 2^-54 of the bound.
// bound == 0 (all contributions zero) gives S = 1; a non-finite bound gives NaN.
__device__ __forceinline__ double fixed_scale_54(double bound) {
  if (!(bound < INFINITY)) return NAN;
  return bound > 0.0 ? ldexp(1.0, 53 - ilogb(bound)) : 1.0;
}

// max that propagates a NaN in either argument
__device__ __forceinline__ double max_nan(double a, double b) { return (a != a || a > b) ? a : b; }

// slot (holding a 64-bit integer) += round(v * S)
__device__ __forceinline__ void lds_add_fixed64(double* slot, double v, double S) {
  atomicAdd(reinterpret_cast<unsigned long long*>(slot), (unsigned long long)llrint(v * S));
}

// Two-word form for sums that need full FP64 resolution: t = v * S is split into its
// nearest integer h (added to hi) and the rest t - h (|t - h| <= 1/2, exact), scaled by
// 2^54 and rounded (added to lo). Each contribution is resolved to 2^-54 / S, about
// 2^-107 of the bound; lo takes up to 1024 contributions per slot without overflow.
__device__ __forceinline__ void lds_add_fixed128(double* hi, double* lo, double v, double S) {
  const double t = v * S;
  const double h = rint(t);
  atomicAdd(reinterpret_cast<unsigned long long*>(hi), (unsigned long long)(long long)h);
  atomicAdd(reinterpret_cast<unsigned long long*>(lo), (unsigned long long)llrint((t - h) * 0x1p54));
}

// the value of a two-word slot pair (inv = 1 / S)
__device__ __forceinline__ double fixed_value2(double hi, double lo, double inv) {
  return (double)__double_as_longlong(hi) * inv + (double)__double_as_longlong(lo) * (inv * 0x1p-54);
}

// the double held as a fixed-point integer in a slot (inv = 1 / S, exact)
__device__ __forceinline__ double fixed_value(double slot, double inv) {
  return (double)__double_as_longlong(slot) * inv;
}

// Maximum of v over the work-group, through red[NT / 64] in LDS: call with every thread,
// then __syncthreads(), then read with block_max_read.
template <int NT>
__device__ __forceinline__ void block_max_write(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max_nan(v, __shfl_down(v, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
}
template <int NT>
__device__ __forceinline__ double block_max_read(const double* red) {
  double m = red[0];
#pragma unroll
  for (int q = 1; q < NT / 64; ++q) m = max_nan(m, red[q]);
  return m;
}

}  // namespace pa
