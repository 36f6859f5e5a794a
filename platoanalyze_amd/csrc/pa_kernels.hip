// Chunked element kernels: one work-group owns a run of consecutive block rows,
// accumulates the contributions of every element touching those rows in LDS
// (atomic-free in HBM: each CRS value and residual entry is written exactly once,
// coalesced), then streams the finished segment out.
#include "pa_kernels.hpp"

namespace pa {

// Work-groups are dispatched round-robin over the 8 XCDs; give each XCD one
// contiguous run of chunks so neighbouring chunks (which share elements and nodes)
// hit the same L2.
__device__ __forceinline__ int xcd_chunk(int b, int n) {
  const int q = n / kNumXcd, r = n % kNumXcd;
  const int x = b % kNumXcd, k = b / kNumXcd;
  return x * q + (x < r ? x : r) + k;
}

__device__ __forceinline__ void lds_add(double* p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <int BITS>
__device__ __forceinline__ int slot_of(const uint32_t* cw, int a, int b) {
  if constexpr (BITS == 4) {
    return (cw[a >> 1] >> (16 * (a & 1) + 4 * b)) & 15u;
  } else {
    return (cw[a] >> (8 * b)) & 255u;
  }
}

// ComputeGradientWorkset<3>: J[d][k] = X_k[d] - X_3[d]; grad N_k = row k of J^-1
// (k<3), grad N_3 = -(sum); vol = |det J| * 1/6 (LinearTetCubRuleDegreeOne weight).
__device__ __forceinline__ void tet_gradients(const double x[4][3], double g[4][3], double& vol) {
  const double j11 = x[0][0] - x[3][0], j12 = x[1][0] - x[3][0], j13 = x[2][0] - x[3][0];
  const double j21 = x[0][1] - x[3][1], j22 = x[1][1] - x[3][1], j23 = x[2][1] - x[3][1];
  const double j31 = x[0][2] - x[3][2], j32 = x[1][2] - x[3][2], j33 = x[2][2] - x[3][2];
  const double det = j11 * j22 * j33 + j12 * j23 * j31 + j13 * j21 * j32 - j11 * j23 * j32 -
                     j12 * j21 * j33 - j13 * j22 * j31;
  const double inv = 1.0 / det;
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
  vol = fabs(det) * (1.0 / 6.0);
}

// Strain<3>: Voigt {xx,yy,zz,yz,xz,xy} with engineering shear.
__device__ __forceinline__ void tet_strain(const double u[4][3], const double g[4][3], double e[6]) {
#pragma unroll
  for (int k = 0; k < 6; ++k) e[k] = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    e[0] += u[a][0] * g[a][0];
    e[1] += u[a][1] * g[a][1];
    e[2] += u[a][2] * g[a][2];
    e[3] += u[a][2] * g[a][1] + u[a][1] * g[a][2];
    e[4] += u[a][2] * g[a][0] + u[a][0] * g[a][2];
    e[5] += u[a][1] * g[a][0] + u[a][0] * g[a][1];
  }
}

// LinearStress<3> with the isotropic Voigt stiffness.
__device__ __forceinline__ void iso_stress(const Material& m, const double e[6], double s[6]) {
  s[0] = m.c00 * e[0] + m.c01 * e[1] + m.c01 * e[2];
  s[1] = m.c01 * e[0] + m.c00 * e[1] + m.c01 * e[2];
  s[2] = m.c01 * e[0] + m.c01 * e[1] + m.c00 * e[2];
  s[3] = m.c33 * e[3];
  s[4] = m.c33 * e[4];
  s[5] = m.c33 * e[5];
}

// StressDivergence<3> row of node a: r_i = vol * sum_j s[V(i,j)] g_a[j].
__device__ __forceinline__ void tet_divergence(const double s[6], const double ga[3], double vol,
                                               double r[3]) {
  r[0] = vol * (s[0] * ga[0] + s[5] * ga[1] + s[4] * ga[2]);
  r[1] = vol * (s[5] * ga[0] + s[1] * ga[1] + s[3] * ga[2]);
  r[2] = vol * (s[4] * ga[0] + s[3] * ga[1] + s[2] * ga[2]);
}

template <int BITS, int MODE>
__global__ __launch_bounds__(kThreads) void k_chunk(ChunkArgs A) {
  extern __shared__ double lds[];
  constexpr bool kJac = MODE == kModeJacobian || MODE == kModeJacobianResidual;
  constexpr bool kRes = MODE == kModeJacobianResidual || MODE == kModeResidual;
  constexpr bool kGzT = MODE == kModeGradZT;
  constexpr bool kEnergy = MODE == kModeEnergy || MODE == kModeEnergyGradZ;
  constexpr bool kGradZ = MODE == kModeEnergyGradZ;
  constexpr bool kNeedU = kRes || kGzT || kEnergy;
  constexpr int kBlkW = kJac ? 9 : (kGzT ? 3 : 0);  // doubles per block in LDS
  constexpr int kRowW = kRes ? 3 : (kGradZ ? 1 : 0);  // doubles per row in LDS

  const int c = xcd_chunk(blockIdx.x, A.nchunks);
  const int r0 = A.chunk_rows[c], r1 = A.chunk_rows[c + 1];
  const int nr = r1 - r0;
  const int b0 = A.rowmap[r0];
  const int nb = A.rowmap[r1] - b0;
  double* seg = lds;
  double* res = lds + (size_t)A.max_blocks * kBlkW;
  int* roff = reinterpret_cast<int*>(res + (size_t)A.max_rows * kRowW);

  if constexpr (kBlkW > 0)
    for (int k = threadIdx.x; k < nb * kBlkW; k += kThreads) seg[k] = 0.0;
  if constexpr (kRowW > 0)
    for (int k = threadIdx.x; k < nr * kRowW; k += kThreads) res[k] = 0.0;
  if constexpr (kBlkW > 0)
    for (int k = threadIdx.x; k <= nr; k += kThreads) roff[k] = A.rowmap[r0 + k] - b0;
  __syncthreads();

  double energy = 0.0;
  const int l0 = A.list_offs[c], l1 = A.list_offs[c + 1];
  for (int k = l0 + threadIdx.x; k < l1; k += kThreads) {
    const uint32_t ent = A.list[k];
    const int e = (int)(ent & 0x0FFFFFFFu);
    const uint32_t mask = ent >> 28;
    const int4 cn = A.conn[e];
    const int nd[4] = {cn.x, cn.y, cn.z, cn.w};
    double x[4][3], g[4][3], vol;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int d = 0; d < 3; ++d) x[a][d] = A.X[3 * (size_t)nd[a] + d];
    tet_gradients(x, g, vol);
    // ApplyWeighting: mean nodal density, MSIMP weight.
    double rho = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) rho += A.z[nd[a]];
    rho = rho / 4.0;
    const double pw = pow(rho, A.m.p);
    const double w = A.m.rmin + (1.0 - A.m.rmin) * pw;

    uint32_t cw[4];
    if constexpr (kBlkW > 0) {
      if constexpr (BITS == 4) {
        const uint2 t = reinterpret_cast<const uint2*>(A.codes)[e];
        cw[0] = t.x;
        cw[1] = t.y;
      } else {
        const uint4 t = reinterpret_cast<const uint4*>(A.codes)[e];
        cw[0] = t.x;
        cw[1] = t.y;
        cw[2] = t.z;
        cw[3] = t.w;
      }
    }

    double s[6], eps[6];
    if constexpr (kNeedU) {
      double u[4][3];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) u[a][d] = A.u[3 * (size_t)nd[a] + d];
      tet_strain(u, g, eps);
      iso_stress(A.m, eps, s);
    }

    if constexpr (kEnergy) {
      // ScalarProduct<6>(sigma_w, eps) * 0.5 * vol ; counted once, by the chunk owning vertex 0.
      double se = 0.0;
#pragma unroll
      for (int q = 0; q < 6; ++q) se += (w * s[q]) * eps[q];
      if (mask & 1u) energy += 0.5 * se * vol;
      if constexpr (kGradZ) {
        // d(0.5 vol sigma_w:eps)/dz_a = 0.5 vol (sigma:eps) * dw/drho * 1/4
        double sr = 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q) sr += s[q] * eps[q];
        const double dw = (1.0 - A.m.rmin) * A.m.p * pow(rho, A.m.p - 1.0);
        const double gz = 0.5 * vol * sr * dw * 0.25;
#pragma unroll
        for (int a = 0; a < 4; ++a)
          if (mask & (1u << a)) lds_add(&res[nd[a] - r0], gz);
      }
    }

    if constexpr (kRes) {
      double sw[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) sw[q] = w * s[q];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        if (mask & (1u << a)) {
          double r[3];
          tet_divergence(sw, g[a], vol, r);
          double* dst = &res[3 * (nd[a] - r0)];
          lds_add(dst + 0, r[0]);
          lds_add(dst + 1, r[1]);
          lds_add(dst + 2, r[2]);
        }
      }
    }

    if constexpr (kJac) {
      const double lam = A.m.c01, mu = A.m.c33;
      const double vw = vol * w;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        if (!(mask & (1u << a))) continue;
        const int base = roff[nd[a] - r0];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int sl = base + slot_of<BITS>(cw, a, b);
          double* dst = &seg[9 * sl];
          const double dot = g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2];
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
              double v = lam * g[a][i] * g[b][j] + mu * g[a][j] * g[b][i];
              if (i == j) v += mu * dot;
              lds_add(dst + 3 * i + j, vw * v);
            }
        }
      }
    }

    if constexpr (kGzT) {
      // dR_b/dz_a = dw/drho * 1/4 * (vol B_b^T sigma), rows = control node a.
      const double dw = (1.0 - A.m.rmin) * A.m.p * pow(rho, A.m.p - 1.0) * 0.25;
      double rb[4][3];
#pragma unroll
      for (int b = 0; b < 4; ++b) tet_divergence(s, g[b], vol, rb[b]);
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        if (!(mask & (1u << a))) continue;
        const int base = roff[nd[a] - r0];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int sl = base + slot_of<BITS>(cw, a, b);
#pragma unroll
          for (int i = 0; i < 3; ++i) lds_add(&seg[3 * sl + i], dw * rb[b][i]);
        }
      }
    }
  }

  if constexpr (kEnergy) {
    // block reduction of the energy (fixed order -> deterministic)
    __shared__ double wsum[kThreads / 64];
    double v = energy;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int q = 0; q < kThreads / 64; ++q) t += wsum[q];
      A.partials[c] = t;
    }
  }
  __syncthreads();
  if constexpr (kRes) {
    if (A.sub_load) {
      for (int k = A.load_offs[c] + threadIdx.x; k < A.load_offs[c + 1]; k += kThreads) {
        const int lr = A.load_row[k];
        res[3 * lr + 0] -= A.load_f[3 * k + 0];
        res[3 * lr + 1] -= A.load_f[3 * k + 1];
        res[3 * lr + 2] -= A.load_f[3 * k + 2];
      }
      __syncthreads();
    }
  }

  if constexpr (kBlkW > 0) {
    double* dst = A.out_vals + (size_t)b0 * kBlkW;
    for (int k = threadIdx.x; k < nb * kBlkW; k += kThreads) dst[k] = seg[k];
  }
  if constexpr (kRowW > 0) {
    double* dst = A.out_vec + (size_t)r0 * kRowW;
    for (int k = threadIdx.x; k < nr * kRowW; k += kThreads) dst[k] = res[k];
  }
}

size_t chunk_lds_bytes(int mode, int max_rows, int max_blocks) {
  int blkw = (mode == kModeJacobian || mode == kModeJacobianResidual) ? 9 : (mode == kModeGradZT ? 3 : 0);
  int roww = (mode == kModeJacobianResidual || mode == kModeResidual) ? 3 : (mode == kModeEnergyGradZ ? 1 : 0);
  size_t bytes = sizeof(double) * ((size_t)max_blocks * blkw + (size_t)max_rows * roww);
  if (blkw > 0) bytes += sizeof(int) * (size_t)(max_rows + 1);
  return bytes;
}

template <int BITS>
static hipError_t launch_chunk_b(int mode, const ChunkArgs& a, hipStream_t s) {
  const size_t lds = chunk_lds_bytes(mode, a.max_rows, a.max_blocks);
  dim3 grid(a.nchunks), block(kThreads);
  switch (mode) {
    case kModeJacobian: k_chunk<BITS, kModeJacobian><<<grid, block, lds, s>>>(a); break;
    case kModeJacobianResidual: k_chunk<BITS, kModeJacobianResidual><<<grid, block, lds, s>>>(a); break;
    case kModeResidual: k_chunk<BITS, kModeResidual><<<grid, block, lds, s>>>(a); break;
    case kModeGradZT: k_chunk<BITS, kModeGradZT><<<grid, block, lds, s>>>(a); break;
    case kModeEnergy: k_chunk<BITS, kModeEnergy><<<grid, block, lds, s>>>(a); break;
    case kModeEnergyGradZ: k_chunk<BITS, kModeEnergyGradZ><<<grid, block, lds, s>>>(a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int BITS, int MODE>
static void allow_lds() {
  hipFuncSetAttribute(reinterpret_cast<const void*>(&k_chunk<BITS, MODE>),
                      hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

hipError_t launch_chunk(int bits, int mode, const ChunkArgs& a, hipStream_t s) {
  static bool once = false;
  if (!once) {
    once = true;
    allow_lds<4, kModeJacobian>();
    allow_lds<4, kModeJacobianResidual>();
    allow_lds<4, kModeGradZT>();
    allow_lds<8, kModeJacobian>();
    allow_lds<8, kModeJacobianResidual>();
    allow_lds<8, kModeGradZT>();
  }
  if (a.nchunks == 0) return hipSuccess;
  return bits == 4 ? launch_chunk_b<4>(mode, a, s) : launch_chunk_b<8>(mode, a, s);
}

}  // namespace pa
