// U-Net kernels: implicit-GEMM 3x3x3 convolutions on v_mfma_f32_16x16x32_bf16, weight
// gradients, packing, Adam and the loss.
//
// MFMA 16x16x32 (bf16 in, fp32 accumulate) operand layout, lane l of 64:
//   A (16 x 32): row l % 16, k = 8 (l / 16) .. 8 (l / 16) + 7
//   B (32 x 16): column l % 16, same k range
//   C (16 x 16): column l % 16, rows 4 (l / 16) .. 4 (l / 16) + 3
#include "pa_unet.hpp"

#include "pa_reduce.hpp"

namespace pa {
namespace unet {

namespace {

typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kT = 256;
constexpr int kConvFillWaves = 2048;  // eight waves per CU of a 256-CU part
constexpr int kConvMaxSlices = 32;

__device__ __forceinline__ bf16_t f2bf(float f) {  // round to nearest even
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7F800000u) == 0x7F800000u) return (bf16_t)(u >> 16);  // inf / nan
  return (bf16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ bool bf_pos(bf16_t b) { return !(b & 0x8000u) && (b & 0x7FFFu); }

// Input voxel linked to output voxel (ox, oy, oz) by tap (dx, dy, dz); false when the
// link falls outside the input grid (zero padding) or off the stride lattice.
__device__ __forceinline__ bool link(const Grid& gi, int s, int transposed, int ox, int oy, int oz, int dx, int dy,
                                     int dz, int& vi) {
  int ix, iy, iz;
  if (!transposed) {
    ix = s * ox + dx - 1;
    iy = s * oy + dy - 1;
    iz = s * oz + dz - 1;
  } else {
    ix = ox + 1 - dx;
    iy = oy + 1 - dy;
    iz = oz + 1 - dz;
    if (ix < 0 || iy < 0 || iz < 0) return false;
    if (s == 2) {
      if ((ix | iy | iz) & 1) return false;
      ix >>= 1;
      iy >>= 1;
      iz >>= 1;
    }
  }
  if (ix < 0 || iy < 0 || iz < 0 || ix >= gi.nx || iy >= gi.ny || iz >= gi.nz) return false;
  vi = (ix * gi.ny + iy) * gi.nz + iz;
  return true;
}

__device__ __forceinline__ void decode(const Grid& g, int o, int& x, int& y, int& z) {
  z = o % g.nz;
  const int r = o / g.nz;
  y = r % g.ny;
  x = r / g.ny;
}

// Bias, ReLU, gate and the stores of output element (m, n) with the reduction sum v.
__device__ __forceinline__ void conv_epilogue(const ConvArgs& A, int m, int n, float v) {
  if (A.bias) v += A.bias[n];
  if (A.relu) v = fmaxf(v, 0.0f);
  if (A.gate && !bf_pos(A.gate[(size_t)m * A.ldg + A.gcol + n])) v = 0.0f;
  if (A.Y) A.Y[(size_t)m * A.ldy + A.ycol + n] = f2bf(v);
  if (A.Yf) {
    float* d = A.Yf + (size_t)m * A.ldyf + A.yfcol + n;
    *d = A.accumulate ? *d + v : v;
  }
}

// One work-group = 4 wavefronts; a wavefront computes 16 output voxels x 64 channels over
// K steps [blockIdx.z * kper, (blockIdx.z + 1) * kper) of 32 (SPLIT: raw sums into the
// slice's part array; else the whole reduction and the epilogue).
template <bool XF32, bool SPLIT>
__global__ __launch_bounds__(kT) void k_conv(ConvArgs A, int kper) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int Vo = A.go.count(), Vi = A.gi.count();
  const int Mtot = A.batch * Vo;
  const int m0 = (blockIdx.x * 4 + wave) * 16;
  const int n0 = blockIdx.y * 64;
  const int Ktot = 27 * A.cin;
  const int kbeg = SPLIT ? (int)blockIdx.z * kper * 32 : 0;
  const int kend = SPLIT ? min(Ktot, kbeg + kper * 32) : Ktot;
  const int mrow = m0 + (lane & 15);
  const bool mvalid = mrow < Mtot;
  int b = 0, ox = 0, oy = 0, oz = 0;
  if (mvalid) {
    b = mrow / Vo;
    decode(A.go, mrow - b * Vo, ox, oy, oz);
  }
  const size_t xbase = (size_t)b * Vi;
  f32x4 acc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
  // this lane's k offset inside a 32-wide K step, tracked as (tap, channel)
  const int kq = 8 * (lane >> 4);
  int t = (kbeg + kq) / A.cin, ch = (kbeg + kq) - ((kbeg + kq) / A.cin) * A.cin;
  for (int kc = kbeg; kc < kend; kc += 32) {
    const int k0 = kc + kq;
    bf16x8 a = {0, 0, 0, 0, 0, 0, 0, 0};
    if (mvalid && k0 < Ktot) {
      int vi;
      if (link(A.gi, A.stride, A.transposed, ox, oy, oz, t / 9, (t / 3) % 3, t % 3, vi)) {
        const size_t off = (xbase + vi) * (size_t)A.ldx + A.xcol + ch;
        if constexpr (XF32) {
          const float4 p = *reinterpret_cast<const float4*>(A.Xf + off);
          const float4 q = *reinterpret_cast<const float4*>(A.Xf + off + 4);
          a = bf16x8{(short)f2bf(p.x), (short)f2bf(p.y), (short)f2bf(p.z), (short)f2bf(p.w),
                     (short)f2bf(q.x), (short)f2bf(q.y), (short)f2bf(q.z), (short)f2bf(q.w)};
        } else {
          a = *reinterpret_cast<const bf16x8*>(A.X + off);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = n0 + 16 * j + (lane & 15);
      bf16x8 bv = {0, 0, 0, 0, 0, 0, 0, 0};
      if (n < A.cout && k0 < Ktot) bv = *reinterpret_cast<const bf16x8*>(A.Bt + (size_t)n * Ktot + k0);
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc[j], 0, 0, 0);
    }
    ch += 32;
    while (ch >= A.cin) {
      ch -= A.cin;
      ++t;
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int n = n0 + 16 * j + (lane & 15);
    if (n >= A.cout) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = m0 + 4 * (lane >> 4) + i;
      if (m >= Mtot) continue;
      if constexpr (SPLIT) A.part[((size_t)blockIdx.z * Mtot + m) * A.cout + n] = acc[j][i];
      else conv_epilogue(A, m, n, acc[j][i]);
    }
  }
}

// Sum of the slices of a split reduction, in slice order, and the epilogue.
__global__ __launch_bounds__(kT) void k_conv_reduce(ConvArgs A, int nslices) {
  const size_t Mtot = (size_t)A.batch * A.go.count(), total = Mtot * A.cout;
  for (size_t e = blockIdx.x * (size_t)kT + threadIdx.x; e < total; e += (size_t)gridDim.x * kT) {
    float v = A.part[e];
    for (int z = 1; z < nslices; ++z) v += A.part[(size_t)z * total + e];
    conv_epilogue(A, (int)(e / A.cout), (int)(e % A.cout), v);
  }
}

// Weight gradient.  Work-group (split, tap x ci-tile, co-tile); each wavefront takes
// 32-voxel K steps of its share of output voxels; fp32 results added atomically.
constexpr int kWgradSplit = 64;

__global__ __launch_bounds__(kT) void k_wgrad(WgradArgs A) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nct = (A.cin + 15) / 16;
  const int t = blockIdx.y / nct, ci0 = (blockIdx.y % nct) * 16;
  const int co0 = blockIdx.z * 64;
  const int dx = t / 9, dy = (t / 3) % 3, dz = t % 3;
  const int Vo = A.go.count(), Vi = A.gi.count();
  const int Mtot = A.batch * Vo;
  const int per = (Mtot + gridDim.x - 1) / gridDim.x;
  const int ms = blockIdx.x * per, me = min(Mtot, ms + per);
  const int ci = ci0 + (lane & 15);
  const int kq = 8 * (lane >> 4);
  f32x4 acc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int mb = ms + wave * 32; mb < me; mb += 4 * 32) {
    bf16x8 a;
    int vin[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int m = mb + kq + q;
      vin[q] = -1;
      if (m < me) {
        const int b = m / Vo;
        int ox, oy, oz, vi;
        decode(A.go, m - b * Vo, ox, oy, oz);
        if (link(A.gi, A.stride, A.transposed, ox, oy, oz, dx, dy, dz, vi)) vin[q] = b * Vi + vi;
      }
      a[q] = (vin[q] >= 0 && ci < A.cin) ? (short)A.X[(size_t)vin[q] * A.ldx + A.xcol + ci] : (short)0;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int co = co0 + 16 * j + (lane & 15);
      bf16x8 bv;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int m = mb + kq + q;
        bv[q] = (m < me && co < A.cout) ? (short)f2bf(A.dY[(size_t)m * A.lddy + A.dycol + co]) : (short)0;
      }
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc[j], 0, 0, 0);
    }
  }
  // C rows = ci (4 (l / 16) + i), columns = co (l % 16)
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int co = co0 + 16 * j + (lane & 15);
    if (co >= A.cout) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int c = ci0 + 4 * (lane >> 4) + i;
      if (c < A.cin) atomicAdd(&A.dW[((size_t)t * A.cin + c) * A.cout + co], acc[j][i]);
    }
  }
}

// db[co] += sum over voxels of dY[m][co]
__global__ __launch_bounds__(kT) void k_bias_grad(const float* dY, int ld, int col, int cout, int M, float* db) {
  const int co = threadIdx.x % 64 + 64 * blockIdx.y;
  const int r = threadIdx.x / 64;
  if (co >= cout) return;
  float s = 0.0f;
  for (int m = blockIdx.x * 4 + r; m < M; m += gridDim.x * 4) s += dY[(size_t)m * ld + col + co];
  atomicAdd(&db[co], s);
}

__global__ __launch_bounds__(kT) void k_pack(const float* W, int cin, int cout, bf16_t* Bf, bf16_t* Bd) {
  const size_t n = (size_t)27 * cin * cout;
  for (size_t e = blockIdx.x * (size_t)kT + threadIdx.x; e < n; e += (size_t)gridDim.x * kT) {
    const int co = (int)(e % cout);
    const size_t r = e / cout;
    const int ci = (int)(r % cin);
    const int t = (int)(r / cin);
    const bf16_t v = f2bf(W[e]);
    if (Bf) Bf[(size_t)co * 27 * cin + (size_t)t * cin + ci] = v;
    if (Bd) Bd[(size_t)ci * 27 * cout + (size_t)t * cout + co] = v;
  }
}

__global__ __launch_bounds__(kT) void k_adam(float* w, const float* g, float* m, float* v, size_t n, float lr,
                                             float b1, float b2, float eps, float c1, float c2) {
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

__global__ __launch_bounds__(kT) void k_mse(const float* p, const float* y, size_t n, float* dp, double* P) {
  __shared__ double sh[kT / 64];
  double s = 0.0;
  const float sc = 2.0f / (float)n;
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) {
    const float d = p[i] - y[i];
    s += (double)d * d;
    dp[i] = sc * d;
  }
  const double t = block_reduce_all<kT>(s, sh);
  if (threadIdx.x == 0) P[blockIdx.x] = t / (double)n;
}

__global__ __launch_bounds__(kT) void k_zero(float* y, size_t n) {
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) y[i] = 0.0f;
}

int blocks_for(size_t n, int cap = 4096) {
  size_t g = (n + kT - 1) / kT;
  return (int)(g < 1 ? 1 : (g > (size_t)cap ? cap : g));
}

}  // namespace

hipError_t launch_conv(const ConvArgs& a, hipStream_t s) {
  const int M = a.batch * a.go.count();
  const int waves = ((M + 15) / 16) * ((a.cout + 63) / 64);
  const int ksteps = (27 * a.cin + 31) / 32;
  // slices of at least two K steps until the grid holds about kConvFillWaves waves
  int nsl = 1;
  if (a.part && waves < kConvFillWaves) {
    nsl = std::min({(kConvFillWaves + waves - 1) / waves, ksteps / 2, kConvMaxSlices});
    while (nsl > 1 && (size_t)nsl * M * a.cout > a.part_cap) --nsl;
  }
  dim3 grid((M + 63) / 64, (a.cout + 63) / 64);
  if (nsl <= 1) {
    if (a.X) k_conv<false, false><<<grid, kT, 0, s>>>(a, ksteps);
    else k_conv<true, false><<<grid, kT, 0, s>>>(a, ksteps);
    return hipGetLastError();
  }
  const int kper = (ksteps + nsl - 1) / nsl;
  nsl = (ksteps + kper - 1) / kper;  // every slice non-empty
  grid.z = nsl;
  if (a.X) k_conv<false, true><<<grid, kT, 0, s>>>(a, kper);
  else k_conv<true, true><<<grid, kT, 0, s>>>(a, kper);
  k_conv_reduce<<<blocks_for((size_t)M * a.cout), kT, 0, s>>>(a, nsl);
  return hipGetLastError();
}

hipError_t launch_wgrad(const WgradArgs& a, hipStream_t s) {
  const int M = a.batch * a.go.count();
  const int split = std::max(1, std::min(kWgradSplit, M / 512));
  dim3 grid(split, 27 * ((a.cin + 15) / 16), (a.cout + 63) / 64);
  k_wgrad<<<grid, kT, 0, s>>>(a);
  if (a.db) {
    dim3 g2(std::max(1, std::min(256, M / 256)), (a.cout + 63) / 64);
    k_bias_grad<<<g2, kT, 0, s>>>(a.dY, a.lddy, a.dycol, a.cout, M, a.db);
  }
  return hipGetLastError();
}

hipError_t launch_pack(const float* W, int cin, int cout, bf16_t* Bfwd, bf16_t* Bdgrad, hipStream_t s) {
  k_pack<<<blocks_for((size_t)27 * cin * cout), kT, 0, s>>>(W, cin, cout, Bfwd, Bdgrad);
  return hipGetLastError();
}

hipError_t launch_adam(float* w, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2,
                       float eps, int t, hipStream_t s) {
  const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
  k_adam<<<blocks_for(n), kT, 0, s>>>(w, g, m, v, n, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t launch_mse(const float* pred, const float* target, size_t n, float* dp, double* partials, int grid,
                      hipStream_t s) {
  k_mse<<<grid, kT, 0, s>>>(pred, target, n, dp, partials);
  return hipGetLastError();
}

hipError_t launch_zero(float* y, size_t n, hipStream_t s) {
  k_zero<<<blocks_for(n), kT, 0, s>>>(y, n);
  return hipGetLastError();
}

}  // namespace unet
}  // namespace pa

// ------------------------------------------------------------------ network (host)
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/plato_mi355x.h"

namespace {

using pa::unet::bf16_t;
using pa::unet::Grid;

constexpr int kInCh = 8;   // density, load (3), coordinates (3), ones
constexpr int kOutCh = 16;  // displacement channels (3, or 9 for three load cases) + padding

// f32 -> bf16 conversion of the network input
__global__ void k_to_bf16(const float* x, bf16_t* y, size_t n) {
  for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const uint32_t u = __float_as_uint(x[i]);
    y[i] = ((u & 0x7F800000u) == 0x7F800000u) ? (bf16_t)(u >> 16) : (bf16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
  }
}

template <typename T>
struct Dev {
  T* p = nullptr;
  size_t n = 0;
  Dev() = default;
  Dev(const Dev&) = delete;
  Dev& operator=(const Dev&) = delete;
  Dev(Dev&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  hipError_t alloc(size_t k) {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = k;
    return k ? hipMalloc(&p, k * sizeof(T)) : hipSuccess;
  }
  ~Dev() {
    if (p) (void)hipFree(p);
  }
};

// A tensor view: buffer, row stride, first channel, channels.
struct View {
  void* p;
  int ld, col, c;
};

struct Layer {
  int cin, cout, stride, transposed, relu;
  int lin, lout;        // grid levels of input and output
  View in, out;         // activations (bf16; out of the last layer is fp32)
  View gin, gout;       // gradients (fp32); gin.p null for the first layer
  size_t w_off, b_off;  // offsets into the parameter vector
  bf16_t *Bf, *Bd;
};

}  // namespace

struct pa_unet_s {
  Grid g[4];
  int base = 16, batch = 0;
  std::vector<Layer> L;
  size_t nparams = 0;
  Dev<float> P, G, M, V;  // parameters, gradients, Adam moments
  Dev<bf16_t> B;          // packed bf16 operands of all layers
  std::vector<Dev<bf16_t>> act;
  std::vector<Dev<float>> grad;
  Dev<float> out, dout;
  Dev<bf16_t> in;
  Dev<double> partials;
  Dev<float> split;  // split-reduction slices of the small convolutions
  double* host_partials = nullptr;
  int step = 0;
  hipStream_t stream = nullptr;
  std::string err;
  ~pa_unet_s() {
    if (host_partials) (void)hipHostFree(host_partials);
  }
};

namespace {

thread_local std::string g_unet_err;

int ufail(pa_unet h, int code, const std::string& msg) {
  if (h) h->err = msg;
  g_unet_err = msg;
  return code;
}

#define PA_UHIP(h, x)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess) return ufail(h, PA_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Buffers per batch size: activations and gradients of every level.
//   act: 0 in(8) [l0], 1 e0(C) [l0], 2 cat0(2C) [l0], 3 e1(2C) [l1], 4 cat1(4C) [l1],
//        5 e2(4C) [l2], 6 cat2(8C) [l2], 7 e3(8C) [l3], 8 b3(8C) [l3], 9 d2(4C) [l2],
//        10 d1(2C) [l1], 11 d0(C) [l0]
struct Buf {
  int level, ch;
};
constexpr int kNumBuf = 12;
constexpr size_t kSplitFloats = (size_t)8 << 20;  // 32 MB of split-reduction slices

void build_layers(pa_unet h) {
  const int C = h->base;
  const Buf bufs[kNumBuf] = {{0, kInCh}, {0, C}, {0, 2 * C}, {1, 2 * C}, {1, 4 * C}, {2, 4 * C},
                             {2, 8 * C}, {3, 8 * C}, {3, 8 * C}, {2, 4 * C}, {1, 2 * C}, {0, C}};
  (void)bufs;
  // (input buffer, col, cin) -> (output buffer, col, cout), stride, transposed, relu
  struct Def {
    int bi, ci, cin, bo, co, cout, s, tr, relu;
  };
  const Def defs[] = {
      {0, 0, kInCh, 1, 0, C, 1, 0, 1},          // enc0
      {1, 0, C, 2, 0, C, 1, 0, 1},              // enc0b -> skip 0
      {2, 0, C, 3, 0, 2 * C, 2, 0, 1},          // down0
      {3, 0, 2 * C, 4, 0, 2 * C, 1, 0, 1},      // enc1 -> skip 1
      {4, 0, 2 * C, 5, 0, 4 * C, 2, 0, 1},      // down1
      {5, 0, 4 * C, 6, 0, 4 * C, 1, 0, 1},      // enc2 -> skip 2
      {6, 0, 4 * C, 7, 0, 8 * C, 2, 0, 1},      // down2
      {7, 0, 8 * C, 8, 0, 8 * C, 1, 0, 1},      // bottleneck
      {8, 0, 8 * C, 6, 4 * C, 4 * C, 2, 1, 1},  // up2 -> cat2
      {6, 0, 8 * C, 9, 0, 4 * C, 1, 0, 1},      // dec2
      {9, 0, 4 * C, 4, 2 * C, 2 * C, 2, 1, 1},  // up1 -> cat1
      {4, 0, 4 * C, 10, 0, 2 * C, 1, 0, 1},     // dec1
      {10, 0, 2 * C, 2, C, C, 2, 1, 1},         // up0 -> cat0
      {2, 0, 2 * C, 11, 0, C, 1, 0, 1},         // dec0
      {11, 0, C, -1, 0, kOutCh, 1, 0, 0},       // output (fp32)
  };
  const int lev[kNumBuf] = {0, 0, 0, 1, 1, 2, 2, 3, 3, 2, 1, 0};
  const int ldb[kNumBuf] = {kInCh, C, 2 * C, 2 * C, 4 * C, 4 * C, 8 * C, 8 * C, 8 * C, 4 * C, 2 * C, C};
  h->L.clear();
  size_t off = 0;
  for (const Def& d : defs) {
    Layer l{};
    l.cin = d.cin;
    l.cout = d.cout;
    l.stride = d.s;
    l.transposed = d.tr;
    l.relu = d.relu;
    l.lin = lev[d.bi];
    l.lout = d.bo >= 0 ? lev[d.bo] : 0;
    l.in = View{(void*)(intptr_t)d.bi, ldb[d.bi], d.ci, d.cin};
    l.out = d.bo >= 0 ? View{(void*)(intptr_t)d.bo, ldb[d.bo], d.co, d.cout} : View{nullptr, kOutCh, 0, kOutCh};
    l.w_off = off;
    off += (size_t)27 * d.cin * d.cout;
    l.b_off = off;
    off += d.cout;
    h->L.push_back(l);
  }
  h->nparams = off;
}

int ensure_batch(pa_unet h, int batch) {
  if (batch == h->batch) return PA_OK;
  const int C = h->base;
  const int lev[kNumBuf] = {0, 0, 0, 1, 1, 2, 2, 3, 3, 2, 1, 0};
  const int ldb[kNumBuf] = {kInCh, C, 2 * C, 2 * C, 4 * C, 4 * C, 8 * C, 8 * C, 8 * C, 4 * C, 2 * C, C};
  h->act.clear();
  h->grad.clear();
  h->act.resize(kNumBuf);
  h->grad.resize(kNumBuf);
  for (int b = 0; b < kNumBuf; ++b) {
    const size_t n = (size_t)batch * h->g[lev[b]].count() * ldb[b];
    PA_UHIP(h, h->act[b].alloc(n));
    PA_UHIP(h, h->grad[b].alloc(n));
  }
  const size_t no = (size_t)batch * h->g[0].count() * kOutCh;
  PA_UHIP(h, h->out.alloc(no));
  PA_UHIP(h, h->dout.alloc(no));
  if (!h->split.p) PA_UHIP(h, h->split.alloc(kSplitFloats));
  h->batch = batch;
  return PA_OK;
}

pa::unet::ConvArgs fwd_args(pa_unet h, const Layer& l) {
  pa::unet::ConvArgs a{};
  const int bi = (int)(intptr_t)l.in.p;
  a.X = h->act[bi].p;
  a.ldx = l.in.ld;
  a.xcol = l.in.col;
  a.cin = l.cin;
  a.gi = h->g[l.lin];
  a.go = h->g[l.lout];
  a.batch = h->batch;
  a.stride = l.stride;
  a.transposed = l.transposed;
  a.Bt = l.Bf;
  a.cout = l.cout;
  a.bias = h->P.p + l.b_off;
  a.relu = l.relu;
  if (l.out.p) {
    a.Y = h->act[(int)(intptr_t)l.out.p].p;
    a.ldy = l.out.ld;
    a.ycol = l.out.col;
  } else {
    a.Yf = h->out.p;
    a.ldyf = kOutCh;
  }
  a.part = h->split.p;
  a.part_cap = h->split.n;
  return a;
}

int forward(pa_unet h) {
  for (const Layer& l : h->L) PA_UHIP(h, pa::unet::launch_conv(fwd_args(h, l), h->stream));
  return PA_OK;
}

int backward(pa_unet h) {
  PA_UHIP(h, pa::unet::launch_zero(h->G.p, h->nparams, h->stream));
  for (int k = (int)h->L.size() - 1; k >= 0; --k) {
    const Layer& l = h->L[k];
    const int bi = (int)(intptr_t)l.in.p;
    // gradient of this layer's output (already multiplied by the ReLU derivative)
    const float* dY = l.out.p ? h->grad[(int)(intptr_t)l.out.p].p : h->dout.p;
    const int lddy = l.out.ld, dycol = l.out.col;
    pa::unet::WgradArgs w{};
    w.X = h->act[bi].p;
    w.ldx = l.in.ld;
    w.xcol = l.in.col;
    w.cin = l.cin;
    w.gi = h->g[l.lin];
    w.go = h->g[l.lout];
    w.batch = h->batch;
    w.stride = l.stride;
    w.transposed = l.transposed;
    w.dY = dY;
    w.lddy = lddy;
    w.dycol = dycol;
    w.cout = l.cout;
    w.dW = h->G.p + l.w_off;
    w.db = h->G.p + l.b_off;
    PA_UHIP(h, pa::unet::launch_wgrad(w, h->stream));
    if (k == 0) continue;  // no gradient for the network input
    pa::unet::ConvArgs a{};
    a.Xf = dY;
    a.ldx = lddy;
    a.xcol = dycol;
    a.cin = l.cout;
    a.gi = h->g[l.lout];
    a.go = h->g[l.lin];
    a.batch = h->batch;
    a.stride = l.stride;
    a.transposed = !l.transposed;
    a.Bt = l.Bd;
    a.cout = l.cin;
    a.Yf = h->grad[bi].p;
    a.ldyf = l.in.ld;
    a.yfcol = l.in.col;
    // skip tensors receive a second gradient from the down-sampling layer (their first
    // consumer in the backward order is the decoder conv that reads the concatenation)
    a.accumulate = (l.stride == 2 && !l.transposed) ? 1 : 0;
    a.gate = h->act[bi].p;  // every layer input except the network input is a ReLU output
    a.ldg = l.in.ld;
    a.gcol = l.in.col;
    a.part = h->split.p;
    a.part_cap = h->split.n;
    PA_UHIP(h, pa::unet::launch_conv(a, h->stream));
  }
  return PA_OK;
}

int repack(pa_unet h) {
  for (const Layer& l : h->L)
    PA_UHIP(h, pa::unet::launch_pack(h->P.p + l.w_off, l.cin, l.cout, l.Bf, l.Bd, h->stream));
  return PA_OK;
}

int load_input(pa_unet h, const float* x, int batch) {
  if (int rc = ensure_batch(h, batch)) return rc;
  const size_t n = (size_t)batch * h->g[0].count() * kInCh;
  k_to_bf16<<<(int)std::min<size_t>(4096, (n + 255) / 256), 256, 0, h->stream>>>(x, h->act[0].p, n);
  PA_UHIP(h, hipGetLastError());
  return PA_OK;
}

int loss_and_grad(pa_unet h, const float* target, double* loss) {
  const size_t n = (size_t)h->batch * h->g[0].count() * kOutCh;
  const int grid = (int)std::min<size_t>(1024, (n + 255) / 256);
  PA_UHIP(h, pa::unet::launch_mse(h->out.p, target, n, h->dout.p, h->partials.p, grid, h->stream));
  PA_UHIP(h, hipMemcpyAsync(h->host_partials, h->partials.p, grid * sizeof(double), hipMemcpyDeviceToHost,
                            h->stream));
  PA_UHIP(h, hipStreamSynchronize(h->stream));
  double s = 0.0;
  for (int i = 0; i < grid; ++i) s += h->host_partials[i];
  if (loss) *loss = s;
  return PA_OK;
}

}  // namespace

extern "C" {

const char* pa_unet_last_error(void) { return g_unet_err.c_str(); }

int pa_unet_create(pa_unet* out, int nx, int ny, int nz, int base_channels, unsigned long long seed) {
  if (!out) return ufail(nullptr, PA_ERR_ARG, "pa_unet_create: null output");
  *out = nullptr;
  if (base_channels < 8 || base_channels % 8) return ufail(nullptr, PA_ERR_ARG, "base_channels must be a multiple of 8");
  if (nx < 5 || ny < 5 || nz < 5 || (nx - 1) % 4 || (ny - 1) % 4 || (nz - 1) % 4)
    return ufail(nullptr, PA_ERR_ARG, "node grid dimensions must be 1 mod 4 and >= 5 (odd on three levels)");
  auto h = new pa_unet_s();
  h->base = base_channels;
  h->g[0] = Grid{nx, ny, nz};
  for (int l = 1; l < 4; ++l)
    h->g[l] = Grid{(h->g[l - 1].nx + 1) / 2, (h->g[l - 1].ny + 1) / 2, (h->g[l - 1].nz + 1) / 2};
  build_layers(h);
  auto cleanup = [&](int rc) {
    delete h;
    return rc;
  };
  if (h->P.alloc(h->nparams) || h->G.alloc(h->nparams) || h->M.alloc(h->nparams) || h->V.alloc(h->nparams) ||
      h->partials.alloc(1024) || hipHostMalloc(&h->host_partials, 1024 * sizeof(double)))
    return cleanup(ufail(nullptr, PA_ERR_HIP, "pa_unet_create: device allocation failed"));
  size_t nb = 0;
  for (const Layer& l : h->L) nb += 2 * (size_t)27 * l.cin * l.cout;
  if (h->B.alloc(nb)) return cleanup(ufail(nullptr, PA_ERR_HIP, "pa_unet_create: device allocation failed"));
  size_t o = 0;
  for (Layer& l : h->L) {
    l.Bf = h->B.p + o;
    o += (size_t)27 * l.cin * l.cout;
    l.Bd = h->B.p + o;
    o += (size_t)27 * l.cin * l.cout;
  }
  // He-normal weights (fan-in 27 cin), zero biases
  std::vector<float> p(h->nparams, 0.0f);
  std::mt19937_64 rng(seed);
  for (const Layer& l : h->L) {
    std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (27.0f * l.cin)));
    for (size_t i = 0; i < (size_t)27 * l.cin * l.cout; ++i) p[l.w_off + i] = nd(rng);
  }
  if (hipMemcpy(h->P.p, p.data(), p.size() * sizeof(float), hipMemcpyHostToDevice) ||
      hipMemset(h->M.p, 0, h->nparams * sizeof(float)) || hipMemset(h->V.p, 0, h->nparams * sizeof(float)))
    return cleanup(ufail(nullptr, PA_ERR_HIP, "pa_unet_create: upload failed"));
  if (repack(h) || hipStreamSynchronize(nullptr)) return cleanup(ufail(nullptr, PA_ERR_HIP, "pa_unet_create: pack failed"));
  *out = h;
  return PA_OK;
}

int pa_unet_destroy(pa_unet h) {
  delete h;
  return PA_OK;
}

int pa_unet_set_stream(pa_unet h, void* stream) {
  if (!h) return ufail(nullptr, PA_ERR_ARG, "null handle");
  h->stream = (hipStream_t)stream;
  return PA_OK;
}

int pa_unet_num_params(pa_unet h, size_t* n) {
  if (!h || !n) return ufail(h, PA_ERR_ARG, "null argument");
  *n = h->nparams;
  return PA_OK;
}

int pa_unet_params(pa_unet h, float* dst, const float* src) {
  if (!h) return ufail(nullptr, PA_ERR_ARG, "null handle");
  if (dst) PA_UHIP(h, hipMemcpyAsync(dst, h->P.p, h->nparams * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
  if (src) {
    PA_UHIP(h, hipMemcpyAsync(h->P.p, src, h->nparams * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
    if (int rc = repack(h)) return rc;
  }
  return PA_OK;
}

int pa_unet_forward(pa_unet h, const float* x, int batch, float* y) {
  if (!h || !x || !y || batch < 1) return ufail(h, PA_ERR_ARG, "pa_unet_forward: bad argument");
  if (int rc = load_input(h, x, batch)) return rc;
  if (int rc = forward(h)) return rc;
  PA_UHIP(h, hipMemcpyAsync(y, h->out.p, (size_t)batch * h->g[0].count() * kOutCh * sizeof(float),
                            hipMemcpyDeviceToDevice, h->stream));
  return PA_OK;
}

int pa_unet_gradient(pa_unet h, const float* x, const float* target, int batch, float* grad, double* loss) {
  if (!h || !x || !target || batch < 1) return ufail(h, PA_ERR_ARG, "pa_unet_gradient: bad argument");
  if (int rc = load_input(h, x, batch)) return rc;
  if (int rc = forward(h)) return rc;
  if (int rc = loss_and_grad(h, target, loss)) return rc;
  if (int rc = backward(h)) return rc;
  if (grad) PA_UHIP(h, hipMemcpyAsync(grad, h->G.p, h->nparams * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
  return PA_OK;
}

int pa_unet_train_step(pa_unet h, const float* x, const float* target, int batch, float lr, double* loss) {
  if (int rc = pa_unet_gradient(h, x, target, batch, nullptr, loss)) return rc;
  ++h->step;
  PA_UHIP(h, pa::unet::launch_adam(h->P.p, h->G.p, h->M.p, h->V.p, h->nparams, lr, 0.9f, 0.999f, 1e-8f, h->step,
                                   h->stream));
  return repack(h);
}

}  // extern "C"
