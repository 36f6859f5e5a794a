// 3-D convolutional U-Net on the node grid of a box mesh (warm start for the solve):
// BF16 activations (channels last), FP32 accumulation on MFMA, FP32 master weights.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

namespace pa {
namespace unet {

using bf16_t = uint16_t;  // raw bfloat16 bits

// Voxel v of a level = (ix * ny + iy) * nz + iz; a batch of B samples stacks B grids.
struct Grid {
  int nx, ny, nz;
  __host__ __device__ int count() const { return nx * ny * nz; }
};

// One 3x3x3 convolution as an implicit GEMM
//   Y[b, o, :] = act( sum_{tap t, input voxel i linked to o by t} X[b, i, :] W_t + bias )
// with the link
//   direct:     i = s*o + d - 1     (per axis, d in {0,1,2}; stride s; zero padding)
//   transposed: o = s*i + d - 1     (i = (o + 1 - d) / s when that is an integer)
// B-operand layout Bt[n][27*K] (n = output channel, K input channels, tap-major rows).
// The same kernel computes data gradients: the gradient of a direct conv is the
// transposed one with Bt built from W_t^T, and vice versa.
struct ConvArgs {
  const bf16_t* X;   // input activations (bf16), row stride ldx
  const float* Xf;   // or FP32 input (converted to bf16 on load) when X == nullptr
  int ldx, xcol, cin;  // row stride, first channel, channels (all multiples of 8)
  Grid gi, go;
  int batch;
  int stride, transposed;
  const bf16_t* Bt;  // [cout][27*cin]
  int cout;
  const float* bias;  // may be null
  int relu;           // apply max(0, .)
  bf16_t* Y;          // bf16 output (may be null), row stride ldy, first channel ycol
  float* Yf;          // fp32 output (may be null), row stride ldyf, first channel yfcol
  int ldy, ycol, ldyf, yfcol;
  int accumulate;     // Yf += result instead of =
  const bf16_t* gate; // multiply the result by (gate > 0) (ReLU derivative), same shape as Yf
  int ldg, gcol;
  // Split reduction for grids too small to fill the GPU: work-group slice z sums its range
  // of the 27*cin reduction into part[z][M][cout] (FP32, M = batch x output voxels) and a
  // second kernel adds the slices in slice order and applies the epilogue.  part null:
  // one pass.  part_cap: floats available at part.
  float* part;
  size_t part_cap;
};
hipError_t launch_conv(const ConvArgs& a, hipStream_t s);

// dW[t][ci][co] += sum over linked (i, o) of X[i][ci] * dY[o][co]; dY fp32 (already
// multiplied by the ReLU derivative).  dW fp32 [27][cin][cout].
struct WgradArgs {
  const bf16_t* X;
  int ldx, xcol, cin;
  Grid gi, go;
  int batch;
  int stride, transposed;
  const float* dY;
  int lddy, dycol, cout;
  float* dW;
  float* db;  // bias gradient (may be null): db[co] += sum_o dY[o][co]
};
hipError_t launch_wgrad(const WgradArgs& a, hipStream_t s);

// Pack FP32 W[27][cin][cout] into bf16 B operands: fwd Bt[co][t][ci] and dgrad
// Bd[ci][t][co] (W_t^T, same tap: the data gradient of a direct conv is the transposed
// link with the same t, and the other way round).
hipError_t launch_pack(const float* W, int cin, int cout, bf16_t* Bfwd, bf16_t* Bdgrad, hipStream_t s);

// Adam step on n parameters (FP32): m, v moments; lr, beta1, beta2, eps, step count t.
hipError_t launch_adam(float* w, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2,
                       float eps, int t, hipStream_t s);

// Mean-squared error over n values: loss partials and gradient 2 (p - y) / n into dp.
hipError_t launch_mse(const float* pred, const float* target, size_t n, float* dp, double* partials, int grid,
                      hipStream_t s);

hipError_t launch_zero(float* y, size_t n, hipStream_t s);

}  // namespace unet
}  // namespace pa
