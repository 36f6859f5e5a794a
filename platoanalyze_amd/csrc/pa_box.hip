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
// and the lane constants are loaded once per work-group).  The loop is software-
// pipelined: while a tile is computed and streamed out, the next tile's gathers and the
// tile after's row data are in flight.
// Work-group barrier for LDS hand-offs only: waits for this wave's LDS operations, not
// for its outstanding global loads (the next tiles' prefetches stay in flight).
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

struct BoxTile {  // a tile's row-level data (first-level loads)
  int base, last, s_base, s_end, s0, gr;
  unsigned long long pm;
};
struct BoxGath {  // a lane's gathered inputs (second-level loads)
  double w[kBoxMaxTerms];
  double ur[3], uc[3];
};

template <bool WITH_R>
__global__ __launch_bounds__(256) void k_box_jac(BoxArgs A) {
  __shared__ double sB[16 * kBoxMaxTerms * 9];
  __shared__ double seg[kBoxRows * kBoxMaxRow * 9];
  for (int q = threadIdx.x; q < 16 * kBoxMaxTerms * 9; q += 256) sB[q] = A.B[q];
  const int lane = threadIdx.x % 16, grp = threadIdx.x / 16;
  // per-lane stencil constants, kept in registers for every tile
  const int cnt = lane == kBoxDiagLane ? 0 : A.S->count[lane];
  int off[kBoxMaxTerms];
#pragma unroll
  for (int e = 0; e < kBoxMaxTerms; ++e) off[e] = A.S->off[lane][e];
  const int d0 = A.S->d[lane][0], d1 = A.S->d[lane][1], d2 = A.S->d[lane][2];
  const int uoff = A.S->uoff[lane];
  const int px = A.nx + 2, py = A.ny + 2;
  const size_t nh = (size_t)px * py * (A.nz + 2);
  __syncthreads();  // block table in LDS
  const double* Bl = sB + lane * (kBoxMaxTerms * 9);
  const int nblk = (A.nrows + kBoxRows - 1) / kBoxRows;
  const int per = (nblk + kNumXcd - 1) / kNumXcd;
  const int x = blockIdx.x % kNumXcd, nbx = gridDim.x / kNumXcd;
  const int t_end = min((x + 1) * per, nblk);
  int blk = x * per + (int)blockIdx.x / kNumXcd;
  if (blk >= t_end) return;  // uniform over the work-group

  auto load1 = [&](int b, BoxTile& T) {
    T.base = b * kBoxRows;
    T.last = min(T.base + kBoxRows, A.nrows);
    T.s_base = A.rowmap[T.base];
    T.s_end = A.rowmap[T.last];
    const int row = T.base + grp;
    T.s0 = 0;
    T.gr = -1;
    T.pm = 0;
    if (row < T.last) {
      T.s0 = A.rowmap[row];
      T.gr = A.gpack[row];
      T.pm = A.perm[row];
    }
  };
  // the lane's stencil neighbour is present when inside the lattice (the host checked
  // that every row holds exactly the present neighbours and itself)
  auto present = [&](const BoxTile& T) {
    const int ri = T.gr & 1023, rj = (T.gr >> 10) & 1023, rk = T.gr >> 20;
    return T.gr >= 0 && cnt > 0 && (unsigned)(ri + d0) <= (unsigned)A.nx && (unsigned)(rj + d1) <= (unsigned)A.ny &&
           (unsigned)(rk + d2) <= (unsigned)A.nz;
  };
  auto lattice = [&](const BoxTile& T) {  // padded index of the row's lattice point
    const int ri = T.gr & 1023, rj = (T.gr >> 10) & 1023, rk = T.gr >> 20;
    return (ri + 1) + px * ((rj + 1) + py * (rk + 1));
  };
  auto load2 = [&](const BoxTile& T, BoxGath& G) {
    const bool nb = present(T);
    const int hr = lattice(T);
#pragma unroll
    for (int e = 0; e < kBoxMaxTerms; ++e) G.w[e] = (nb && e < cnt) ? A.W[hr + off[e]] : 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      G.ur[i] = 0.0;
      G.uc[i] = 0.0;
    }
    if constexpr (WITH_R) {
      if (nb) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          G.ur[i] = A.ulat[i * nh + hr];
          G.uc[i] = A.ulat[i * nh + hr + uoff];
        }
      }
    }
  };

  BoxTile T0, T1;
  BoxGath G0;
  load1(blk, T0);
  load2(T0, G0);
  bool has1 = blk + nbx < t_end;
  if (has1) load1(blk + nbx, T1);
  for (; blk < t_end; blk += nbx) {
    // T0/G0: this tile; T1: the next tile's rows (when has1)
    BoxTile T2;
    BoxGath G1;
    if (has1) load2(T1, G1);
    const bool has2 = blk + 2 * nbx < t_end;
    if (has2) load1(blk + 2 * nbx, T2);

    const int row = T0.base + grp;
    const bool live = row < T0.last;
    const bool nb = present(T0);
    const bool diag = live && lane == kBoxDiagLane;
    const int slot = (int)((T0.pm >> (4 * lane)) & 15);
    double Kv[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) Kv[q] = 0.0;
    if (nb) {
#pragma unroll
      for (int e = 0; e < kBoxMaxTerms; ++e)
#pragma unroll
        for (int q = 0; q < 9; ++q) Kv[q] = fma(G0.w[e], Bl[9 * e + q], Kv[q]);
    }
    double S[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) S[q] = sum16(Kv[q]);
    if (diag) {
#pragma unroll
      for (int q = 0; q < 9; ++q) Kv[q] = -S[q];
    }
    if constexpr (WITH_R) {
      double du[3], ru[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) du[i] = nb ? G0.uc[i] - G0.ur[i] : 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i) ru[i] = sum16(Kv[3 * i] * du[0] + Kv[3 * i + 1] * du[1] + Kv[3 * i + 2] * du[2]);
      if (live && lane < 3) {
        const int k = A.lidx[row];
        const double f = k >= 0 ? A.load_f[3 * (size_t)k + lane] : 0.0;
        A.R[3 * (size_t)row + lane] = (lane == 0 ? ru[0] : (lane == 1 ? ru[1] : ru[2])) - f;
      }
    }
    if (nb || diag) {
      double* d = seg + 9 * (T0.s0 - T0.s_base + slot);
#pragma unroll
      for (int q = 0; q < 9; ++q) d[q] = Kv[q];
    }
    lds_barrier();
    double* out = A.vals + 9 * (size_t)T0.s_base;
    for (int q = threadIdx.x; q < 9 * (T0.s_end - T0.s_base); q += 256) __builtin_nontemporal_store(seg[q], out + q);
    lds_barrier();  // seg is rewritten by the next tile
    T0 = T1;
    G0 = G1;
    T1 = T2;
    has1 = has2;
  }
}

template __global__ void k_box_jac<false>(BoxArgs);
template __global__ void k_box_jac<true>(BoxArgs);

}  // namespace pa
