// Cross-lane sums inside aligned groups of 16 lanes (one DPP row), without LDS.
#pragma once
#include <hip/hip_runtime.h>

namespace pa {

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  // every source lane of these patterns is valid, so no "old" value is needed
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
// DPP controls: quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror, row_mirror
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140;

// Sum over each aligned 16-lane group; every lane of the group receives the same bits.
// After the two quad steps each lane holds its quad's sum; the half-mirror pairs each
// lane with one of the other quad of its half, the mirror with one of the other half.
__device__ __forceinline__ double sum16(double v) {
  v += dpp_f64<kDppXor1>(v);
  v += dpp_f64<kDppXor2>(v);
  v += dpp_f64<kDppHalfMirror>(v);
  v += dpp_f64<kDppMirror>(v);
  return v;
}

}  // namespace pa
