// Block-CRS (3x3) row product shared by the SpMV and the fused multigrid kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace pa {

constexpr int kLanesPerRow = 16;  // 4 block rows per wavefront

// 16 lanes per block row; lane l takes the row's blocks l, l+16, ... (9 contiguous
// values, one column index, three gathered x values each), so every lane issues all
// of its loads back to back.  Row sums are reduced across the 16 lanes.
// X(c, j) returns the j-th component of the input at node c.  V: double, or float for
// an FP32 copy of the matrix (products and sums in FP64).
template <typename XF, typename V = double>
__device__ __forceinline__ void spmv_row(const V* __restrict__ vals, const int* __restrict__ rowmap,
                                         const int* __restrict__ col, int row, int lane, const XF& X,
                                         double acc[3]) {
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  const int s0 = rowmap[row], s1 = rowmap[row + 1];
  for (int s = s0 + lane; s < s1; s += kLanesPerRow) {
    const int c = col[s];
    const V* v = vals + 9 * (size_t)s;
    const double x0 = X(c, 0), x1 = X(c, 1), x2 = X(c, 2);
    a0 += (double)v[0] * x0 + (double)v[1] * x1 + (double)v[2] * x2;
    a1 += (double)v[3] * x0 + (double)v[4] * x1 + (double)v[5] * x2;
    a2 += (double)v[6] * x0 + (double)v[7] * x1 + (double)v[8] * x2;
  }
#pragma unroll
  for (int off = kLanesPerRow / 2; off > 0; off >>= 1) {
    a0 += __shfl_xor(a0, off, kLanesPerRow);
    a1 += __shfl_xor(a1, off, kLanesPerRow);
    a2 += __shfl_xor(a2, off, kLanesPerRow);
  }
  acc[0] = a0;
  acc[1] = a1;
  acc[2] = a2;
}

}  // namespace pa
