// Lattice Jacobian kernels (pa_box.hpp).
#include "pa_box.hpp"
#include "pa_dpp.hpp"
#include "pa_reduce.hpp"

namespace pa {

__global__ __launch_bounds__(256) void k_box_material(const double* P, double* B, double lam, double mu) {
  for (int q = threadIdx.x; q < 16 * kBoxMaxTerms * 9; q += 256) {
    const int ij = q % 9, i = ij / 3, j = ij % 3;
    const double* p = P + (q - ij);
    const double tr = p[0] + p[4] + p[8];
    B[q] = lam * p[ij] + mu * (p[3 * j + i] + (i == j ? tr : 0.0));
  }
}

__global__ __launch_bounds__(256) void k_box_scatter(BoxArgs A) {
  const int px = A.nx + 2, py = A.ny + 2;
  const size_t nh = (size_t)px * py * (A.nz + 2);
  for (int n = blockIdx.x * 256 + threadIdx.x; n < A.nrows; n += gridDim.x * 256) {
    const int g = A.gpack[n];
    const size_t hp = (size_t)((g & 1023) + 1) + (size_t)px * ((((g >> 10) & 1023) + 1) + (size_t)py * ((g >> 20) + 1));
    A.zlat[hp] = A.z[n];
    if (A.ulat) {
#pragma unroll
      for (int c = 0; c < 3; ++c) A.ulat[c * nh + hp] = A.u[3 * (size_t)n + c];
    }
  }
}

__global__ __launch_bounds__(256) void k_box_w(BoxArgs A) {
  const int px = A.nx + 2, py = A.ny + 2, pz = A.nz + 2;
  const size_t nh = (size_t)px * py * pz;
  for (size_t hp = blockIdx.x * (size_t)256 + threadIdx.x; hp < nh; hp += (size_t)gridDim.x * 256) {
    const int I = (int)(hp % px), J = (int)((hp / px) % py), K = (int)(hp / ((size_t)px * py));
    if (I < 1 || J < 1 || K < 1 || I > A.nx || J > A.ny || K > A.nz) {  // ghost hex
#pragma unroll
      for (int t = 0; t < 6; ++t) A.W[t * nh + hp] = 0.0;
      continue;
    }
    double zc[8];  // corners of the hex whose lower corner is lattice point (I-1, J-1, K-1)
#pragma unroll
    for (int c = 0; c < 8; ++c) zc[c] = A.zlat[hp + (c & 1) + px * (((c >> 1) & 1) + (size_t)py * (c >> 2))];
#pragma unroll
    for (int t = 0; t < 6; ++t) {
      const uint8_t* cr = A.corners + 4 * t;
      const double rho = (zc[cr[0]] + zc[cr[1]] + zc[cr[2]] + zc[cr[3]]) / 4.0;
      A.W[t * nh + hp] = penalty_w(A.m, rho);
    }
  }
}

// 16 lanes per graph row.  Lane l of every row computes the block of stencil neighbour
// l (lane 15 the diagonal), so all rows read the same block table entries at the same
// lane: the table reads are LDS broadcasts without bank conflicts, and a lane's element
// offsets stay in registers.  Off-diagonal blocks sum at most six element terms,
// gathered together; the diagonal block is -(sum of the row's other blocks) (every
// element block row sums to zero: sum_b grad N_b = 0); R_r = sum_{c != r} K_rc (u_c -
// u_r) - f_r; 16-lane sums by DPP.  The CRS slot of each lane's block comes from the
// row's slot word.  A work-group's 16 consecutive rows are one contiguous CRS segment,
// staged in LDS and streamed out with coalesced non-temporal stores.  Work-groups are
// dealt to the XCDs in contiguous row ranges (round-robin dispatch puts group b on XCD
// b % 8) and looped over within each XCD's range (a persistent grid: the block table
// and the lane constants are loaded once per work-group). 
// Work-group barrier for LDS hand-offs only: waits for this wave's LDS operations, not
// for its outstanding global stores.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Sums of 12 values over each aligned 16-lane group, scattered over the lanes (a
// reduce-scatter in four DPP exchanges): lane i receives the group sum of value index
// rs_index(i), or nothing when rs_index(i) < 0.  Each exchange pairs every lane with a
// partner holding the complementary half of the current value set.
__device__ __forceinline__ int rs_index(int lane) {
  const int o = ((lane & 2) ? 2 : 0) + (lane & 1);  // offset within the last three values
  return o > 2 ? -1 : (lane < 8 ? 0 : 6) + ((lane & 7) < 4 ? 0 : 3) + o;
}
__device__ __forceinline__ double reduce_scatter12(const double v[12], int lane) {
  const bool l1 = lane < 8, l2 = (lane & 7) < 4, l3 = (lane & 2) == 0, l4 = (lane & 1) == 0;
  double a[6], b[3];
#pragma unroll
  for (int k = 0; k < 6; ++k) a[k] = (l1 ? v[k] : v[6 + k]) + dpp_f64<kDppMirror>(l1 ? v[6 + k] : v[k]);
#pragma unroll
  for (int k = 0; k < 3; ++k) b[k] = (l2 ? a[k] : a[3 + k]) + dpp_f64<kDppHalfMirror>(l2 ? a[3 + k] : a[k]);
  const double c0 = (l3 ? b[0] : b[2]) + dpp_f64<kDppXor2>(l3 ? b[2] : b[0]);
  const double c1 = (l3 ? b[1] : 0.0) + dpp_f64<kDppXor2>(l3 ? 0.0 : b[1]);
  return (l4 ? c0 : c1) + dpp_f64<kDppXor1>(l4 ? c1 : c0);
}

template <bool WITH_R>
__global__ __launch_bounds__(256) void k_box_jac(BoxArgs A) {
  __shared__ double sB[16 * kBoxMaxTerms * 9];
  __shared__ double seg[kBoxRows * kBoxMaxRow * 9];
  for (int q = threadIdx.x; q < 16 * kBoxMaxTerms * 9; q += 256) sB[q] = A.B[q];
  const int lane = threadIdx.x % 16, grp = threadIdx.x / 16;
  // per-lane stencil constants, kept in registers for every tile
  const int cnt = lane == kBoxDiagLane ? 0 : A.S->count[lane];
  unsigned off[kBoxMaxTerms];
#pragma unroll
  for (int e = 0; e < kBoxMaxTerms; ++e) off[e] = (unsigned)A.S->off[lane][e];
  const int d0 = A.S->d[lane][0], d1 = A.S->d[lane][1], d2 = A.S->d[lane][2];
  const unsigned uoff = (unsigned)A.S->uoff[lane];
  const unsigned px = A.nx + 2, py = A.ny + 2;
  const unsigned nh = px * py * (unsigned)(A.nz + 2);
  const double* __restrict__ W = A.W;
  const double* __restrict__ ul = A.ulat;
  __syncthreads();  // block table in LDS
  const double* Bl = sB + lane * (kBoxMaxTerms * 9);
  const int nblk = (A.nrows + kBoxRows - 1) / kBoxRows;
  const int per = (nblk + kNumXcd - 1) / kNumXcd;
  const int x = blockIdx.x % kNumXcd, nbx = gridDim.x / kNumXcd;
  const int t_end = min((x + 1) * per, nblk);
  for (int blk = x * per + (int)blockIdx.x / kNumXcd; blk < t_end; blk += nbx) {
    const int base = blk * kBoxRows;
    const int row = base + grp;
    const int last = min(base + kBoxRows, A.nrows);
    const int s_base = A.rowmap[base], s_end = A.rowmap[last];
    const bool live = row < last;
    int s0 = 0, gr = 0;
    unsigned long long pm = 0;
    if (live) {
      s0 = A.rowmap[row];
      gr = A.gpack[row];
      pm = A.perm[row];
    }
    const int ri = gr & 1023, rj = (gr >> 10) & 1023, rk = gr >> 20;
    const int slot = (int)((pm >> (4 * lane)) & 15);
    // this lane's stencil neighbour: present when inside the lattice (the host checked
    // that every row holds exactly the present neighbours and itself)
    const bool nb = live && cnt > 0 && (unsigned)(ri + d0) <= (unsigned)A.nx && (unsigned)(rj + d1) <= (unsigned)A.ny &&
                    (unsigned)(rk + d2) <= (unsigned)A.nz;
    const unsigned hr = (unsigned)(ri + 1) + px * ((unsigned)(rj + 1) + py * (unsigned)(rk + 1));
    double Kv[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) Kv[q] = 0.0;
    double du[3] = {0.0, 0.0, 0.0};
    if (nb) {
      double w[kBoxMaxTerms];
#pragma unroll
      for (int e = 0; e < kBoxMaxTerms; ++e) w[e] = e < cnt ? W[hr + off[e]] : 0.0;
      if constexpr (WITH_R) {
#pragma unroll
        for (int i = 0; i < 3; ++i) du[i] = ul[i * nh + hr + uoff] - ul[i * nh + hr];
      }
#pragma unroll
      for (int e = 0; e < kBoxMaxTerms; ++e)
#pragma unroll
        for (int q = 0; q < 9; ++q) Kv[q] = fma(w[e], Bl[9 * e + q], Kv[q]);
    }
    // group sums: the nine entries of the diagonal block (-sum of the row's other blocks)
    // and the three residual components, each landing on one lane
    double v12[12];
#pragma unroll
    for (int q = 0; q < 9; ++q) v12[q] = Kv[q];
#pragma unroll
    for (int i = 0; i < 3; ++i) v12[9 + i] = Kv[3 * i] * du[0] + Kv[3 * i + 1] * du[1] + Kv[3 * i + 2] * du[2];
    const double sum = reduce_scatter12(v12, lane);
    const int idx = rs_index(lane);
    if (live && idx >= 0 && idx < 9) {
      const int dslot = (int)((pm >> (4 * kBoxDiagLane)) & 15);
      seg[9 * (s0 - s_base + dslot) + idx] = -sum;
    }
    if constexpr (WITH_R) {
      if (live && idx >= 9) {
        const int k = A.lidx[row];
        const double f = k >= 0 ? A.load_f[3 * (size_t)k + (idx - 9)] : 0.0;
        A.R[3 * (size_t)row + (idx - 9)] = sum - f;
      }
    }
    if (nb) {
      double* d = seg + 9 * (s0 - s_base + slot);
#pragma unroll
      for (int q = 0; q < 9; ++q) d[q] = Kv[q];
    }
    lds_barrier();
    {  // 16-byte stores from the first 16-byte-aligned value on
      typedef double v2d __attribute__((ext_vector_type(2)));
      double* out = A.vals + 9 * (size_t)s_base;
      const int n = 9 * (s_end - s_base);
      const int head = s_base & 1;  // 9 s_base odd: the segment starts 8 bytes past a 16-byte boundary
      if (head && threadIdx.x == 0) __builtin_nontemporal_store(seg[0], out);
      for (int q = head + 2 * (int)threadIdx.x; q + 1 < n; q += 512)
        __builtin_nontemporal_store(v2d{seg[q], seg[q + 1]}, reinterpret_cast<v2d*>(out + q));
      if (((n - head) & 1) && threadIdx.x == 0) __builtin_nontemporal_store(seg[n - 1], out + n - 1);
    }
    lds_barrier();  // seg is rewritten by the next tile
  }
}

template __global__ void k_box_jac<false>(BoxArgs);
template __global__ void k_box_jac<true>(BoxArgs);

}  // namespace pa
