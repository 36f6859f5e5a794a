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

__global__ __launch_bounds__(256) void k_box_w(BoxArgs A) {
  const int px = A.nx + 2, py = A.ny + 2, pz = A.nz + 2;
  const size_t nh = (size_t)px * py * pz;
  for (size_t hp = blockIdx.x * (size_t)256 + threadIdx.x; hp < nh; hp += (size_t)gridDim.x * 256) {
    const int I = (int)(hp % px), J = (int)((hp / px) % py), K = (int)(hp / ((size_t)px * py));
    const int i = I - 1, j = J - 1, k = K - 1;
    if (i < 0 || j < 0 || k < 0 || i >= A.nx || j >= A.ny || k >= A.nz) {
#pragma unroll
      for (int t = 0; t < 6; ++t) A.W[t * nh + hp] = 0.0;
      continue;
    }
    double zc[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int ci = i + (c & 1), cj = j + ((c >> 1) & 1), ck = k + (c >> 2);
      zc[c] = A.z[A.nog[ci + (A.nx + 1) * (cj + (size_t)(A.ny + 1) * ck)]];
    }
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
// b % 8).
template <bool WITH_R>
__global__ __launch_bounds__(256) void k_box_jac(BoxArgs A) {
  __shared__ double sB[16 * kBoxMaxTerms * 9];
  __shared__ double seg[kBoxRows * kBoxMaxRow * 9];
  const int nblk = (A.nrows + kBoxRows - 1) / kBoxRows;
  const int per = (nblk + kNumXcd - 1) / kNumXcd;
  const int blk = (blockIdx.x % kNumXcd) * per + blockIdx.x / kNumXcd;
  if (blk >= nblk) return;  // uniform over the work-group
  for (int q = threadIdx.x; q < 16 * kBoxMaxTerms * 9; q += 256) sB[q] = A.B[q];
  const int lane = threadIdx.x % 16, grp = threadIdx.x / 16;
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
  const int cnt = lane == kBoxDiagLane ? 0 : A.S->count[lane];
  const bool nb = live && cnt > 0 && (unsigned)(ri + A.S->d[lane][0]) <= (unsigned)A.nx &&
                  (unsigned)(rj + A.S->d[lane][1]) <= (unsigned)A.ny &&
                  (unsigned)(rk + A.S->d[lane][2]) <= (unsigned)A.nz;
  const bool diag = live && lane == kBoxDiagLane;
  double w[kBoxMaxTerms];
#pragma unroll
  for (int e = 0; e < kBoxMaxTerms; ++e) w[e] = 0.0;
  if (nb) {
    const int px = A.nx + 2, py = A.ny + 2;
    const int hp = (ri + 1) + px * ((rj + 1) + py * (rk + 1));  // padded hex with lower corner r
#pragma unroll
    for (int e = 0; e < kBoxMaxTerms; ++e)
      if (e < cnt) w[e] = A.W[hp + A.S->off[lane][e]];
  }
  double du[3] = {0.0, 0.0, 0.0};
  if constexpr (WITH_R) {
    if (nb) {
      const int c = A.col[s0 + slot];
#pragma unroll
      for (int i = 0; i < 3; ++i) du[i] = A.u[3 * (size_t)c + i] - A.u[3 * (size_t)row + i];
    }
  }
  __syncthreads();  // block table in LDS
  double Kv[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) Kv[q] = 0.0;
  if (nb) {
    const double* Bl = sB + lane * (kBoxMaxTerms * 9);
#pragma unroll
    for (int e = 0; e < kBoxMaxTerms; ++e)
#pragma unroll
      for (int q = 0; q < 9; ++q) Kv[q] = fma(w[e], Bl[9 * e + q], Kv[q]);
  }
  double S[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) S[q] = sum16(Kv[q]);
  if (diag) {
#pragma unroll
    for (int q = 0; q < 9; ++q) Kv[q] = -S[q];
  }
  if constexpr (WITH_R) {
    double ru[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) ru[i] = sum16(Kv[3 * i] * du[0] + Kv[3 * i + 1] * du[1] + Kv[3 * i + 2] * du[2]);
    if (live && lane < 3) {
      const int k = A.lidx[row];
      const double f = k >= 0 ? A.load_f[3 * (size_t)k + lane] : 0.0;
      A.R[3 * (size_t)row + lane] = (lane == 0 ? ru[0] : (lane == 1 ? ru[1] : ru[2])) - f;
    }
  }
  if (nb || diag) {
    double* d = seg + 9 * (s0 - s_base + slot);
#pragma unroll
    for (int q = 0; q < 9; ++q) d[q] = Kv[q];
  }
  __syncthreads();
  double* out = A.vals + 9 * (size_t)s_base;
  for (int q = threadIdx.x; q < 9 * (s_end - s_base); q += 256) __builtin_nontemporal_store(seg[q], out + q);
}

template __global__ void k_box_jac<false>(BoxArgs);
template __global__ void k_box_jac<true>(BoxArgs);

}  // namespace pa
