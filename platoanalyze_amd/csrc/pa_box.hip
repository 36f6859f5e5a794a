// Lattice Jacobian kernels (pa_box.hpp).
#include "pa_box.hpp"
#include "pa_dpp.hpp"
#include "pa_reduce.hpp"
#include "pa_solver.hpp"

namespace pa {

__global__ __launch_bounds__(256) void k_box_material(const double* P, double* B, float* Bf, double lam,
                                                      double mu) {
  for (int q = threadIdx.x; q < 16 * kBoxMaxTerms * 9; q += 256) {
    const int ij = q % 9, i = ij / 3, j = ij % 3;
    const double* p = P + (q - ij);
    const double tr = p[0] + p[4] + p[8];
    B[q] = lam * p[ij] + mu * (p[3 * j + i] + (i == j ? tr : 0.0));
    Bf[q] = (float)B[q];
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
      for (int c = 0; c < 3; ++c) A.ulat[3 * hp + c] = A.u[3 * (size_t)n + c];
    }
  }
}

template <bool SIMP_INT, typename T>
__global__ __launch_bounds__(256) void k_box_w(BoxArgs A, T* W) {
  // One wave per run of 64 padded hexes along x: the run's line (J, K) and start are
  // wave-uniform, so the index arithmetic is 32-bit scalar work (6 nh < 2^32 was checked
  // at set-up).
  const unsigned px = A.nx + 2, py = A.ny + 2, pz = A.nz + 2;
  const unsigned nseg = (px + 63) / 64, nwave = nseg * py * pz;
  const unsigned lane = threadIdx.x % 64;
  const unsigned w0 = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + threadIdx.x / 64);
  int cr[24];  // corner of each local vertex of the six tets, wave-uniform (scalar registers)
#pragma unroll
  for (int q = 0; q < 24; ++q) cr[q] = __builtin_amdgcn_readfirstlane((int)A.corners[q]);
  for (unsigned w = w0; w < nwave; w += gridDim.x * 4) {
    const unsigned line = w / nseg, I = (w - line * nseg) * 64 + lane;
    const unsigned J = line % py, K = line / py;
    if (I >= px) continue;
    const unsigned hp = I + px * line;
    typedef T v2 __attribute__((ext_vector_type(2)));
    v2* w2 = reinterpret_cast<v2*>(W + 6 * (size_t)hp);  // 6-weight records: three 2-vector stores
    if (I < 1 || J < 1 || K < 1 || I > (unsigned)A.nx || J > (unsigned)A.ny || K > (unsigned)A.nz) {  // ghost hex
#pragma unroll
      for (int k = 0; k < 3; ++k) w2[k] = v2{T(0), T(0)};
      continue;
    }
    double zc[8];  // corners of the hex whose lower corner is lattice point (I-1, J-1, K-1)
#pragma unroll
    for (int c = 0; c < 8; ++c) zc[c] = A.zlat[hp + (c & 1) + px * (((c >> 1) & 1) + py * (c >> 2))];
    double w6[6];
#pragma unroll
    for (int t = 0; t < 6; ++t) {
      const double rho = (zc[cr[4 * t]] + zc[cr[4 * t + 1]] + zc[cr[4 * t + 2]] + zc[cr[4 * t + 3]]) / 4.0;
      // SIMP with a small integer exponent: penalty_w's own multiplication path, inlined
      // without the other models' code
      w6[t] = SIMP_INT ? A.m.rmin + (1.0 - A.m.rmin) * simp_pow(rho, A.m.p, A.m.pint) : penalty_w(A.m, rho);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) w2[k] = v2{(T)w6[2 * k], (T)w6[2 * k + 1]};
  }
}
template __global__ void k_box_w<true, double>(BoxArgs, double*);
template __global__ void k_box_w<false, double>(BoxArgs, double*);
template __global__ void k_box_w<true, float>(BoxArgs, float*);
template __global__ void k_box_w<false, float>(BoxArgs, float*);

// 8 lanes per graph row; lane j computes the blocks of stencil neighbours j and j + 8
// (stencil lane 15 is the diagonal), so all rows read the same block table entries at
// the same lane: the table reads are LDS broadcasts without bank conflicts, and a lane's
// element offsets stay in registers.  An off-diagonal block sums at most six element
// terms; the diagonal block is -(sum of the row's other blocks) (every element block row
// sums to zero: sum_b grad N_b = 0); R_r = sum_{c != r} K_rc (u_c - u_r) - f_r.  The
// nine diagonal entries and three residual components are summed over the row's lanes
// in one DPP reduce-scatter.  The CRS slot of each block comes from the row's slot
// word.  A work-group's 32 consecutive rows are one contiguous CRS segment, staged in LDS
// in two halves of 16 rows and streamed out with coalesced non-temporal stores.  Work-groups are dealt to the
// XCDs in contiguous row ranges (round-robin dispatch puts group b on XCD b % 8) and loop
// over tiles within their XCD's range (a persistent grid: the block table and the lane
// constants are loaded once per work-group).

// Work-group barrier for LDS hand-offs only: waits for this wave's LDS operations, not
// for its outstanding global stores.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Sums of 12 values over each aligned 8-lane group, scattered: lane j (0..7) receives
// the group sums of value indices rs_first(j) (in s0) and, for even j, rs_first(j) + 1
// (in s1).  Three DPP exchanges, each pairing a lane with the partner holding the
// complementary half of the current value set.
__device__ __forceinline__ int rs_first(int j) { return (j < 4 ? 0 : 6) + ((j & 2) ? 3 : 0) + ((j & 1) ? 2 : 0); }
__device__ __forceinline__ void reduce_scatter12x8(const double v[12], int j, double& s0, double& s1) {
  const bool l1 = j < 4, l2 = (j & 2) == 0, l3 = (j & 1) == 0;
  double a[6], b[3];
#pragma unroll
  for (int k = 0; k < 6; ++k) a[k] = (l1 ? v[k] : v[6 + k]) + dpp_f64<kDppHalfMirror>(l1 ? v[6 + k] : v[k]);
#pragma unroll
  for (int k = 0; k < 3; ++k) b[k] = (l2 ? a[k] : a[3 + k]) + dpp_f64<kDppXor2>(l2 ? a[3 + k] : a[k]);
  s0 = (l3 ? b[0] : b[2]) + dpp_f64<kDppXor1>(l3 ? b[2] : b[0]);
  s1 = (l3 ? b[1] : 0.0) + dpp_f64<kDppXor1>(l3 ? 0.0 : b[1]);
}

template <bool WITH_R>
__global__ __launch_bounds__(256, 4) void k_box_jac(BoxArgs A) {
  __shared__ double sB[16 * kBoxMaxTerms * 9];
  __shared__ double seg[(kBoxRows / 2) * kBoxMaxRow * 9];  // one half-tile of 16 rows
  for (int q = threadIdx.x; q < 16 * kBoxMaxTerms * 9; q += 256) sB[q] = A.B[q];
  const int j = threadIdx.x % 8, grp = threadIdx.x / 8;
  // constants of the lane's two stencil neighbours (j and j + 8), kept in registers
  const int la = j, lb = j + 8;
  const int cnta = A.S->count[la], cntb = lb == kBoxDiagLane ? 0 : A.S->count[lb];
  unsigned offa[kBoxMaxTerms], offb[kBoxMaxTerms];
#pragma unroll
  for (int e = 0; e < kBoxMaxTerms; ++e) {
    offa[e] = (unsigned)A.S->off[la][e];
    offb[e] = (unsigned)A.S->off[lb][e];
  }
  const int da0 = A.S->d[la][0], da1 = A.S->d[la][1], da2 = A.S->d[la][2];
  const int db0 = A.S->d[lb][0], db1 = A.S->d[lb][1], db2 = A.S->d[lb][2];
  const unsigned uoffa = (unsigned)A.S->uoff[la], uoffb = (unsigned)A.S->uoff[lb];
  const unsigned px = A.nx + 2, py = A.ny + 2;
  const unsigned nh = px * py * (unsigned)(A.nz + 2);
  const double* __restrict__ W = A.W;
  const double* __restrict__ ul = A.ulat;
  __syncthreads();  // block table in LDS
  const double* Ba = sB + la * (kBoxMaxTerms * 9);
  const double* Bb = sB + lb * (kBoxMaxTerms * 9);
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
    // a stencil neighbour is present when inside the lattice (the host checked that
    // every row holds exactly the present neighbours and itself)
    const bool nba = live && cnta > 0 && (unsigned)(ri + da0) <= (unsigned)A.nx &&
                     (unsigned)(rj + da1) <= (unsigned)A.ny && (unsigned)(rk + da2) <= (unsigned)A.nz;
    const bool nbb = live && cntb > 0 && (unsigned)(ri + db0) <= (unsigned)A.nx &&
                     (unsigned)(rj + db1) <= (unsigned)A.ny && (unsigned)(rk + db2) <= (unsigned)A.nz;
    const unsigned hr = (unsigned)(ri + 1) + px * ((unsigned)(rj + 1) + py * (unsigned)(rk + 1));
    // Every index below stays inside the padded arrays for every lane: terms a lane does
    // not have point at offset 0 with a zero block, ghost hexes weigh zero, and a lane
    // without a neighbour reads the row's own u (du = 0).
    double wa[kBoxMaxTerms], wb[kBoxMaxTerms];
#pragma unroll
    for (int e = 0; e < kBoxMaxTerms; ++e) {
      wa[e] = W[6 * hr + offa[e]];
      wb[e] = W[6 * hr + offb[e]];
    }
    double dua[3] = {0.0, 0.0, 0.0}, dub[3] = {0.0, 0.0, 0.0};
    if constexpr (WITH_R) {
      const unsigned hca = nba ? hr + uoffa : hr, hcb = nbb ? hr + uoffb : hr;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const double ur = ul[3 * hr + i];
        dua[i] = ul[3 * hca + i] - ur;
        dub[i] = ul[3 * hcb + i] - ur;
      }
    }
    double Ka[9], Kb[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      Ka[q] = wa[0] * Ba[q];
      Kb[q] = wb[0] * Bb[q];
    }
#pragma unroll
    for (int e = 1; e < kBoxMaxTerms; ++e)
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        Ka[q] = fma(wa[e], Ba[9 * e + q], Ka[q]);
        Kb[q] = fma(wb[e], Bb[9 * e + q], Kb[q]);
      }
    // row sums over the 8 lanes: the diagonal block (-sum of the other blocks) and R
    double v12[12];
#pragma unroll
    for (int q = 0; q < 9; ++q) v12[q] = Ka[q] + Kb[q];
#pragma unroll
    for (int i = 0; i < 3; ++i)
      v12[9 + i] = Ka[3 * i] * dua[0] + Ka[3 * i + 1] * dua[1] + Ka[3 * i + 2] * dua[2] +
                   Kb[3 * i] * dub[0] + Kb[3 * i + 1] * dub[1] + Kb[3 * i + 2] * dub[2];
    double sum0, sum1;
    reduce_scatter12x8(v12, j, sum0, sum1);
    if (live) {
      const int i0 = rs_first(j);
      if constexpr (WITH_R) {
        const int k = A.lidx[row];
        if (i0 >= 9) {
          const double f = k >= 0 ? A.load_f[3 * (size_t)k + (i0 - 9)] : 0.0;
          A.R[3 * (size_t)row + (i0 - 9)] = sum0 - f;
        }
        if ((j & 1) == 0 && i0 + 1 >= 9) {
          const double f = k >= 0 ? A.load_f[3 * (size_t)k + (i0 - 8)] : 0.0;
          A.R[3 * (size_t)row + (i0 - 8)] = sum1 - f;
        }
      }
    }
    // the tile's two halves of 16 rows are staged and streamed out in turn
    const int mid = min(base + kBoxRows / 2, last);
    const int s_mid = A.rowmap[mid];
#pragma unroll
    for (int hf = 0; hf < 2; ++hf) {
      const int sb = hf ? s_mid : s_base, se = hf ? s_end : s_mid;
      if ((grp >> 4) == hf) {
        if (live) {
          const int i0 = rs_first(j);
          double* dg = seg + 9 * (s0 - sb + (int)((pm >> (4 * kBoxDiagLane)) & 15));
          if (i0 < 9) dg[i0] = -sum0;
          if ((j & 1) == 0 && i0 + 1 < 9) dg[i0 + 1] = -sum1;
        }
        if (nba) {
          double* d = seg + 9 * (s0 - sb + (int)((pm >> (4 * la)) & 15));
#pragma unroll
          for (int q = 0; q < 9; ++q) d[q] = Ka[q];
        }
        if (nbb) {
          double* d = seg + 9 * (s0 - sb + (int)((pm >> (4 * lb)) & 15));
#pragma unroll
          for (int q = 0; q < 9; ++q) d[q] = Kb[q];
        }
      }
      lds_barrier();
      {  // 16-byte stores from the first 16-byte-aligned value on
        typedef double v2d __attribute__((ext_vector_type(2)));
        double* out = A.vals + 9 * (size_t)sb;
        const int n = 9 * (se - sb);
        const int head = sb & 1;  // 9 sb odd: the segment starts 8 bytes past a 16-byte boundary
        if (head && threadIdx.x == 0 && n > 0) __builtin_nontemporal_store(seg[0], out);
        for (int q = head + 2 * (int)threadIdx.x; q + 1 < n; q += 512)
          __builtin_nontemporal_store(v2d{seg[q], seg[q + 1]}, reinterpret_cast<v2d*>(out + q));
        if (n > head && ((n - head) & 1) && threadIdx.x == 0) __builtin_nontemporal_store(seg[n - 1], out + n - 1);
      }
      lds_barrier();  // seg is rewritten by the next half
    }
  }
}

template __global__ void k_box_jac<false>(BoxArgs);
template __global__ void k_box_jac<true>(BoxArgs);

template <bool WITH_R>
__global__ __launch_bounds__(kBoxWsThreads, 2) void k_box_jac_ws(BoxArgs A) {
  // LDS: two tile buffers of CRS values, two of residual rows, the block table, and the
  // CRS range [s_base, s_end) and first row of the tile in each buffer.
  extern __shared__ double dyn[];
  __shared__ double sB[16 * kBoxMaxTerms * 9];
  __shared__ int hdr[2][3];
  const int cap = A.seg_cap;
  for (int q = threadIdx.x; q < 16 * kBoxMaxTerms * 9; q += kBoxWsThreads) sB[q] = A.B[q];
  const bool producer = threadIdx.x < 256;
  const int tid = producer ? (int)threadIdx.x : (int)threadIdx.x - 256;
  const int j = tid % 8, grp = tid / 8;
  const int la = j, lb = j + 8;
  const int cnta = A.S->count[la], cntb = lb == kBoxDiagLane ? 0 : A.S->count[lb];
  unsigned offa[kBoxMaxTerms], offb[kBoxMaxTerms];
#pragma unroll
  for (int e = 0; e < kBoxMaxTerms; ++e) {
    offa[e] = (unsigned)A.S->off[la][e];
    offb[e] = (unsigned)A.S->off[lb][e];
  }
  const int da0 = A.S->d[la][0], da1 = A.S->d[la][1], da2 = A.S->d[la][2];
  const int db0 = A.S->d[lb][0], db1 = A.S->d[lb][1], db2 = A.S->d[lb][2];
  const unsigned uoffa = (unsigned)A.S->uoff[la], uoffb = (unsigned)A.S->uoff[lb];
  const unsigned px = A.nx + 2, py = A.ny + 2;
  const double* __restrict__ W = A.W;
  const double* __restrict__ ul = A.ulat;
  const double* Ba = sB + la * (kBoxMaxTerms * 9);
  const double* Bb = sB + lb * (kBoxMaxTerms * 9);
  const int nblk = (A.nrows + kBoxRows - 1) / kBoxRows;
  const int per = (nblk + kNumXcd - 1) / kNumXcd;
  const int x = blockIdx.x % kNumXcd, nbx = gridDim.x / kNumXcd;
  const int t0 = x * per + (int)blockIdx.x / kNumXcd, t_end = min((x + 1) * per, nblk);
  const int ntile = t0 < t_end ? (t_end - t0 + nbx - 1) / nbx : 0;
  // row data of the producer's next tile, loaded one tile ahead (the gathers of a tile
  // then wait on one dependent load only)
  int n_sb = 0, n_se = 0, n_s0 = 0, n_gr = 0, n_kl = -1;
  unsigned long long n_pm = 0;
  auto load_rows = [&](int kk) {
    const int base = (t0 + kk * nbx) * kBoxRows;
    const int row = base + grp;
    const int last = min(base + kBoxRows, A.nrows);
    n_sb = A.rowmap[base];
    n_se = A.rowmap[last];
    if (row < last) {
      n_s0 = A.rowmap[row];
      n_gr = A.gpack[row];
      n_pm = A.perm[row];
      if constexpr (WITH_R) n_kl = A.lidx[row];
    }
  };
  if (producer && ntile > 0) load_rows(0);
  __syncthreads();  // block table in LDS
  // iteration k: the producers compute tile k into buffer k & 1 while the consumers
  // stream tile k - 1 out of the other buffer; one barrier per iteration hands over
  for (int k = 0; k <= ntile; ++k) {
    const int b = k & 1;
    if (producer) {
      if (k < ntile) {
        double* seg = dyn + (size_t)b * 9 * cap;
        double* rseg = dyn + (size_t)2 * 9 * cap + b * 3 * kBoxRows;
        const int base = (t0 + k * nbx) * kBoxRows;
        const int row = base + grp;
        const int last = min(base + kBoxRows, A.nrows);
        const bool live = row < last;
        const int s_base = n_sb, s0 = n_s0, gr = live ? n_gr : 0, kl = n_kl;
        const unsigned long long pm = n_pm;
        if (tid == 0) {
          hdr[b][0] = s_base;
          hdr[b][1] = n_se;
          hdr[b][2] = base;
        }
        if (k + 1 < ntile) load_rows(k + 1);
        const int ri = gr & 1023, rj = (gr >> 10) & 1023, rk = gr >> 20;
        const bool nba = live && cnta > 0 && (unsigned)(ri + da0) <= (unsigned)A.nx &&
                         (unsigned)(rj + da1) <= (unsigned)A.ny && (unsigned)(rk + da2) <= (unsigned)A.nz;
        const bool nbb = live && cntb > 0 && (unsigned)(ri + db0) <= (unsigned)A.nx &&
                         (unsigned)(rj + db1) <= (unsigned)A.ny && (unsigned)(rk + db2) <= (unsigned)A.nz;
        const unsigned hr = (unsigned)(ri + 1) + px * ((unsigned)(rj + 1) + py * (unsigned)(rk + 1));
        // in-bounds gathers for every lane, as in k_box_jac
        double wa[kBoxMaxTerms], wb[kBoxMaxTerms];
#pragma unroll
        for (int e = 0; e < kBoxMaxTerms; ++e) {
          wa[e] = W[6 * hr + offa[e]];
          wb[e] = W[6 * hr + offb[e]];
        }
        double dua[3] = {0.0, 0.0, 0.0}, dub[3] = {0.0, 0.0, 0.0};
        if constexpr (WITH_R) {
          const unsigned hca = nba ? hr + uoffa : hr, hcb = nbb ? hr + uoffb : hr;
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            const double ur = ul[3 * hr + i];
            dua[i] = ul[3 * hca + i] - ur;
            dub[i] = ul[3 * hcb + i] - ur;
          }
        }
        double Ka[9], Kb[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
          Ka[q] = wa[0] * Ba[q];
          Kb[q] = wb[0] * Bb[q];
        }
#pragma unroll
        for (int e = 1; e < kBoxMaxTerms; ++e)
#pragma unroll
          for (int q = 0; q < 9; ++q) {
            Ka[q] = fma(wa[e], Ba[9 * e + q], Ka[q]);
            Kb[q] = fma(wb[e], Bb[9 * e + q], Kb[q]);
          }
        double v12[12];
#pragma unroll
        for (int q = 0; q < 9; ++q) v12[q] = Ka[q] + Kb[q];
#pragma unroll
        for (int i = 0; i < 3; ++i)
          v12[9 + i] = Ka[3 * i] * dua[0] + Ka[3 * i + 1] * dua[1] + Ka[3 * i + 2] * dua[2] +
                       Kb[3 * i] * dub[0] + Kb[3 * i + 1] * dub[1] + Kb[3 * i + 2] * dub[2];
        double sum0, sum1;
        reduce_scatter12x8(v12, j, sum0, sum1);
        if (live) {
          const int i0 = rs_first(j);
          double* dg = seg + 9 * (s0 - s_base + (int)((pm >> (4 * kBoxDiagLane)) & 15));
          if (i0 < 9) dg[i0] = -sum0;
          if ((j & 1) == 0 && i0 + 1 < 9) dg[i0 + 1] = -sum1;
          if constexpr (WITH_R) {
            if (i0 >= 9) rseg[3 * grp + i0 - 9] = sum0 - (kl >= 0 ? A.load_f[3 * (size_t)kl + (i0 - 9)] : 0.0);
            if ((j & 1) == 0 && i0 + 1 >= 9)
              rseg[3 * grp + i0 - 8] = sum1 - (kl >= 0 ? A.load_f[3 * (size_t)kl + (i0 - 8)] : 0.0);
          }
        }
        if (nba) {
          double* d = seg + 9 * (s0 - s_base + (int)((pm >> (4 * la)) & 15));
#pragma unroll
          for (int q = 0; q < 9; ++q) d[q] = Ka[q];
        }
        if (nbb) {
          double* d = seg + 9 * (s0 - s_base + (int)((pm >> (4 * lb)) & 15));
#pragma unroll
          for (int q = 0; q < 9; ++q) d[q] = Kb[q];
        }
      }
    } else if (k > 0) {
      // stream tile k - 1: CRS values with 16-byte non-temporal stores from the first
      // 16-byte-aligned value on, then its residual rows (base is a multiple of 32, so
      // 3 base doubles start on a 16-byte boundary)
      typedef double v2d __attribute__((ext_vector_type(2)));
      const double* seg = dyn + (size_t)(b ^ 1) * 9 * cap;
      const int sb = hdr[b ^ 1][0], se = hdr[b ^ 1][1], base = hdr[b ^ 1][2];
      double* out = A.vals + 9 * (size_t)sb;
      const int n = 9 * (se - sb);
      const int head = sb & 1;
      if (head && tid == 0 && n > 0) __builtin_nontemporal_store(seg[0], out);
      for (int q = head + 2 * tid; q + 1 < n; q += 2 * (kBoxWsThreads - 256))
        __builtin_nontemporal_store(v2d{seg[q], seg[q + 1]}, reinterpret_cast<v2d*>(out + q));
      if (n > head && ((n - head) & 1) && tid == 0) __builtin_nontemporal_store(seg[n - 1], out + n - 1);
      if constexpr (WITH_R) {
        const double* rseg = dyn + (size_t)2 * 9 * cap + (b ^ 1) * 3 * kBoxRows;
        const int nr = 3 * (min(base + kBoxRows, A.nrows) - base);
        double* rout = A.R + 3 * (size_t)base;
        if (2 * tid + 1 < nr)
          __builtin_nontemporal_store(v2d{rseg[2 * tid], rseg[2 * tid + 1]}, reinterpret_cast<v2d*>(rout + 2 * tid));
        else if (2 * tid < nr)
          __builtin_nontemporal_store(rseg[2 * tid], rout + 2 * tid);
      }
    }
    lds_barrier();
  }
}
template __global__ void k_box_jac_ws<false>(BoxArgs);
template __global__ void k_box_jac_ws<true>(BoxArgs);

template <bool WITH_R>
__global__ __launch_bounds__(kBoxWsThreads, 2) void k_box_jac_ws_ahead(BoxArgs A) {
  // LDS: two tile buffers of CRS values, two of residual rows, the block table, and the
  // CRS range [s_base, s_end) and first row of the tile in each buffer.
  extern __shared__ double dyn[];
  __shared__ double sB[16 * kBoxMaxTerms * 9];
  __shared__ int hdr[2][3];
  const int cap = A.seg_cap;
  for (int q = threadIdx.x; q < 16 * kBoxMaxTerms * 9; q += kBoxWsThreads) sB[q] = A.B[q];
  // four computing waves (threads 0..255, 8 lanes per row of a 32-row tile) and two
  // streaming waves (threads 256.., ctid 0..127)
  const bool producer = threadIdx.x < 256;
  const int tid = threadIdx.x & 255, ctid = (int)threadIdx.x - 256;
  const int j = tid % 8, grp = tid / 8;
  const int la = j, lb = j + 8;
  const int cnta = A.S->count[la], cntb = lb == kBoxDiagLane ? 0 : A.S->count[lb];
  unsigned offa[kBoxMaxTerms], offb[kBoxMaxTerms];
#pragma unroll
  for (int e = 0; e < kBoxMaxTerms; ++e) {
    offa[e] = (unsigned)A.S->off[la][e];
    offb[e] = (unsigned)A.S->off[lb][e];
  }
  const int da0 = A.S->d[la][0], da1 = A.S->d[la][1], da2 = A.S->d[la][2];
  const int db0 = A.S->d[lb][0], db1 = A.S->d[lb][1], db2 = A.S->d[lb][2];
  const unsigned uoffa = (unsigned)A.S->uoff[la], uoffb = (unsigned)A.S->uoff[lb];
  const unsigned px = A.nx + 2, py = A.ny + 2;
  const double* __restrict__ W = A.W;
  const double* __restrict__ ul = A.ulat;
  const double* Ba = sB + la * (kBoxMaxTerms * 9);
  const double* Bb = sB + lb * (kBoxMaxTerms * 9);
  const int nblk = (A.nrows + kBoxRows - 1) / kBoxRows;
  const int per = (nblk + kNumXcd - 1) / kNumXcd;
  const int x = blockIdx.x % kNumXcd, nbx = gridDim.x / kNumXcd;
  const int t0 = x * per + (int)blockIdx.x / kNumXcd, t_end = min((x + 1) * per, nblk);
  const int ntile = t0 < t_end ? (t_end - t0 + nbx - 1) / nbx : 0;
  const int i0 = rs_first(j);  // first of this lane's two entries of the row sums
  // The computing waves' loads are all unconditional (clamped rows, selects instead of
  // branches), so the compiler's count of loads issued after a tile's gathers is exact:
  // a tile's gathers are issued right after the previous tile's arithmetic and are
  // waited for one staging and one hand-over later.
  struct Rows {
    int sb, se, s0, gr;
    unsigned long long pm;
  };
  struct Gath {
    double wa[kBoxMaxTerms], wb[kBoxMaxTerms], ur[3], ua[3], ub[3];
  };
  // row data of tile kk (clamped to the last tile; a row past the end of its tile reads
  // the tile's last row and is masked where it matters)
  auto load_rows = [&](int kk, Rows& r) {
    const int base = (t0 + min(kk, ntile - 1) * nbx) * kBoxRows;
    const int last = min(base + kBoxRows, A.nrows);
    const int row = min(base + grp, last - 1);
    r.sb = A.rowmap[base];
    r.se = A.rowmap[last];
    r.s0 = A.rowmap[row];
    r.gr = A.gpack[row];
    r.pm = A.perm[row];
  };
  // in-bounds gathers for every lane: terms a lane does not have point at offset 0 with a
  // zero block, ghost hexes weigh zero, a lane without a neighbour reads the row's own u
  auto gather = [&](const Rows& r, Gath& g) {
    const int ri = r.gr & 1023, rj = (r.gr >> 10) & 1023, rk = r.gr >> 20;
    const bool ina = cnta > 0 && (unsigned)(ri + da0) <= (unsigned)A.nx && (unsigned)(rj + da1) <= (unsigned)A.ny &&
                     (unsigned)(rk + da2) <= (unsigned)A.nz;
    const bool inb = cntb > 0 && (unsigned)(ri + db0) <= (unsigned)A.nx && (unsigned)(rj + db1) <= (unsigned)A.ny &&
                     (unsigned)(rk + db2) <= (unsigned)A.nz;
    const unsigned hr = (unsigned)(ri + 1) + px * ((unsigned)(rj + 1) + py * (unsigned)(rk + 1));
#pragma unroll
    for (int e = 0; e < kBoxMaxTerms; ++e) {
      g.wa[e] = W[6 * hr + offa[e]];
      g.wb[e] = W[6 * hr + offb[e]];
    }
    if constexpr (WITH_R) {
      const unsigned hca = ina ? hr + uoffa : hr, hcb = inb ? hr + uoffb : hr;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        g.ur[i] = ul[3 * hr + i];
        g.ua[i] = ul[3 * hca + i];
        g.ub[i] = ul[3 * hcb + i];
      }
    }
  };
  // row data of tiles k, k + 1 and k + 2 rotate through three sets (the loop below is
  // unrolled by three), so no loaded register is copied, which would wait for its load
  Rows rA{}, rB{}, rC{};
  Gath g{};
  if (producer && ntile > 0) {
    load_rows(0, rA);
    load_rows(1, rB);
    gather(rA, g);
  }
  __syncthreads();  // block table in LDS
  // iteration k: the producers compute tile k into buffer k & 1 while the consumers
  // stream tile k - 1 out of the other buffer; one barrier per iteration hands over
  auto iteration = [&](int k, const Rows& rc, const Rows& rn, Rows& r2) __attribute__((always_inline)) {
    const int b = k & 1;
    if (producer) {
      if (k < ntile) {
        double* seg = dyn + (size_t)b * 9 * cap;
        double* rseg = dyn + (size_t)2 * 9 * cap + b * 3 * kBoxRows;
        const int base = (t0 + k * nbx) * kBoxRows;
        const bool live = base + grp < min(base + kBoxRows, A.nrows);
        const int ri = rc.gr & 1023, rj = (rc.gr >> 10) & 1023, rk = rc.gr >> 20;
        const bool nba = live && cnta > 0 && (unsigned)(ri + da0) <= (unsigned)A.nx &&
                         (unsigned)(rj + da1) <= (unsigned)A.ny && (unsigned)(rk + da2) <= (unsigned)A.nz;
        const bool nbb = live && cntb > 0 && (unsigned)(ri + db0) <= (unsigned)A.nx &&
                         (unsigned)(rj + db1) <= (unsigned)A.ny && (unsigned)(rk + db2) <= (unsigned)A.nz;
        double dua[3] = {0.0, 0.0, 0.0}, dub[3] = {0.0, 0.0, 0.0};
        if constexpr (WITH_R) {
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            dua[i] = g.ua[i] - g.ur[i];
            dub[i] = g.ub[i] - g.ur[i];
          }
        }
        double Ka[9], Kb[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
          Ka[q] = g.wa[0] * Ba[q];
          Kb[q] = g.wb[0] * Bb[q];
        }
#pragma unroll
        for (int e = 1; e < kBoxMaxTerms; ++e)
#pragma unroll
          for (int q = 0; q < 9; ++q) {
            Ka[q] = fma(g.wa[e], Ba[9 * e + q], Ka[q]);
            Kb[q] = fma(g.wb[e], Bb[9 * e + q], Kb[q]);
          }
        double v12[12];
#pragma unroll
        for (int q = 0; q < 9; ++q) v12[q] = Ka[q] + Kb[q];
#pragma unroll
        for (int i = 0; i < 3; ++i)
          v12[9 + i] = Ka[3 * i] * dua[0] + Ka[3 * i + 1] * dua[1] + Ka[3 * i + 2] * dua[2] +
                       Kb[3 * i] * dub[0] + Kb[3 * i + 1] * dub[1] + Kb[3 * i + 2] * dub[2];
        double sum0, sum1;
        reduce_scatter12x8(v12, j, sum0, sum1);
        // the next tiles' loads go out now: row data two tiles ahead, gathers one ahead
        load_rows(k + 2, r2);
        gather(rn, g);
        if (tid == 0) {
          hdr[b][0] = rc.sb;
          hdr[b][1] = rc.se;
          hdr[b][2] = base;
        }
        if (live) {
          double* dg = seg + 9 * (rc.s0 - rc.sb + (int)((rc.pm >> (4 * kBoxDiagLane)) & 15));
          if (i0 < 9) dg[i0] = -sum0;
          if ((j & 1) == 0 && i0 + 1 < 9) dg[i0 + 1] = -sum1;
          if constexpr (WITH_R) {
            if (i0 >= 9) rseg[3 * grp + i0 - 9] = sum0;
            if ((j & 1) == 0 && i0 + 1 >= 9) rseg[3 * grp + i0 - 8] = sum1;
          }
        }
        if (nba) {
          double* d = seg + 9 * (rc.s0 - rc.sb + (int)((rc.pm >> (4 * la)) & 15));
#pragma unroll
          for (int q = 0; q < 9; ++q) d[q] = Ka[q];
        }
        if (nbb) {
          double* d = seg + 9 * (rc.s0 - rc.sb + (int)((rc.pm >> (4 * lb)) & 15));
#pragma unroll
          for (int q = 0; q < 9; ++q) d[q] = Kb[q];
        }
      }
    } else if (k > 0) {
      // stream tile k - 1: CRS values with 16-byte non-temporal stores from the first
      // 16-byte-aligned value on, then its residual rows (base is a multiple of 32, so
      // 3 base doubles start on a 16-byte boundary)
      typedef double v2d __attribute__((ext_vector_type(2)));
      const double* seg = dyn + (size_t)(b ^ 1) * 9 * cap;
      const int sb = hdr[b ^ 1][0], se = hdr[b ^ 1][1], base = hdr[b ^ 1][2];
      double* out = A.vals + 9 * (size_t)sb;
      const int n = 9 * (se - sb);
      const int head = sb & 1;
      constexpr int kCons = kBoxWsThreads - 256, kU = 4;
      if (head && ctid == 0 && n > 0) __builtin_nontemporal_store(seg[0], out);
      // kU pairs per thread read from LDS back to back, then stored
      for (int q0 = head + 2 * ctid; q0 + 1 < n; q0 += 2 * kCons * kU) {
        v2d v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          const int q = q0 + 2 * kCons * u;
          if (q + 1 < n) v[u] = v2d{seg[q], seg[q + 1]};
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          const int q = q0 + 2 * kCons * u;
          if (q + 1 < n) __builtin_nontemporal_store(v[u], reinterpret_cast<v2d*>(out + q));
        }
      }
      if (n > head && ((n - head) & 1) && ctid == 0) __builtin_nontemporal_store(seg[n - 1], out + n - 1);
      if constexpr (WITH_R) {
        static_assert(2 * kCons >= 3 * kBoxRows, "one residual pair per streaming lane");
        const double* rseg = dyn + (size_t)2 * 9 * cap + (b ^ 1) * 3 * kBoxRows;
        const int nr = 3 * (min(base + kBoxRows, A.nrows) - base);
        double* rout = A.R + 3 * (size_t)base;
        if (2 * ctid + 1 < nr)
          __builtin_nontemporal_store(v2d{rseg[2 * ctid], rseg[2 * ctid + 1]},
                                      reinterpret_cast<v2d*>(rout + 2 * ctid));
        else if (2 * ctid < nr)
          __builtin_nontemporal_store(rseg[2 * ctid], rout + 2 * ctid);
      }
    }
    lds_barrier();
  };
  for (int k = 0; k <= ntile; k += 3) {
    iteration(k, rA, rB, rC);
    if (k + 1 <= ntile) iteration(k + 1, rB, rC, rA);
    if (k + 2 <= ntile) iteration(k + 2, rC, rA, rB);
  }
}
template __global__ void k_box_jac_ws_ahead<false>(BoxArgs);
template __global__ void k_box_jac_ws_ahead<true>(BoxArgs);

__global__ __launch_bounds__(256, 5) void k_box_res(BoxArgs A, int sub_load) {
  __shared__ double sB[16 * kBoxMaxTerms * 9];
  for (int q = threadIdx.x; q < 16 * kBoxMaxTerms * 9; q += 256) sB[q] = A.B[q];
  const int j = threadIdx.x % 8, grp = threadIdx.x / 8;
  const int la = j, lb = j + 8;
  const int cnta = A.S->count[la], cntb = lb == kBoxDiagLane ? 0 : A.S->count[lb];
  unsigned offa[kBoxMaxTerms], offb[kBoxMaxTerms];
#pragma unroll
  for (int e = 0; e < kBoxMaxTerms; ++e) {
    offa[e] = (unsigned)A.S->off[la][e];
    offb[e] = (unsigned)A.S->off[lb][e];
  }
  const int da0 = A.S->d[la][0], da1 = A.S->d[la][1], da2 = A.S->d[la][2];
  const int db0 = A.S->d[lb][0], db1 = A.S->d[lb][1], db2 = A.S->d[lb][2];
  const unsigned uoffa = (unsigned)A.S->uoff[la], uoffb = (unsigned)A.S->uoff[lb];
  const unsigned px = A.nx + 2, py = A.ny + 2;
  const double* __restrict__ W = A.W;
  const double* __restrict__ ul = A.ulat;
  __syncthreads();  // block table in LDS
  const double* Ba = sB + la * (kBoxMaxTerms * 9);
  const double* Bb = sB + lb * (kBoxMaxTerms * 9);
  const int nblk = (A.nrows + kBoxRows - 1) / kBoxRows;
  const int per = (nblk + kNumXcd - 1) / kNumXcd;
  const int x = blockIdx.x % kNumXcd, nbx = gridDim.x / kNumXcd;
  const int t_end = min((x + 1) * per, nblk);
#pragma nounroll
  for (int blk = x * per + (int)blockIdx.x / kNumXcd; blk < t_end; blk += nbx) {
    // the block table is read from LDS per row, not held in registers across tiles
    asm volatile("" ::: "memory");
    const int row = blk * kBoxRows + grp;
    const bool live = row < A.nrows;
    const int gr = live ? A.gpack[row] : 0;
    const int ri = gr & 1023, rj = (gr >> 10) & 1023, rk = gr >> 20;
    const bool nba = live && cnta > 0 && (unsigned)(ri + da0) <= (unsigned)A.nx &&
                     (unsigned)(rj + da1) <= (unsigned)A.ny && (unsigned)(rk + da2) <= (unsigned)A.nz;
    const bool nbb = live && cntb > 0 && (unsigned)(ri + db0) <= (unsigned)A.nx &&
                     (unsigned)(rj + db1) <= (unsigned)A.ny && (unsigned)(rk + db2) <= (unsigned)A.nz;
    const unsigned hr = (unsigned)(ri + 1) + px * ((unsigned)(rj + 1) + py * (unsigned)(rk + 1));
    double wa[kBoxMaxTerms], wb[kBoxMaxTerms];
#pragma unroll
    for (int e = 0; e < kBoxMaxTerms; ++e) {
      wa[e] = W[6 * hr + offa[e]];
      wb[e] = W[6 * hr + offb[e]];
    }
    const unsigned hca = nba ? hr + uoffa : hr, hcb = nbb ? hr + uoffb : hr;
    double dua[3], dub[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double ur = ul[3 * hr + i];
      dua[i] = ul[3 * hca + i] - ur;
      dub[i] = ul[3 * hcb + i] - ur;
    }
    double Ka[9], Kb[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      Ka[q] = wa[0] * Ba[q];
      Kb[q] = wb[0] * Bb[q];
    }
#pragma unroll
    for (int e = 1; e < kBoxMaxTerms; ++e)
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        Ka[q] = fma(wa[e], Ba[9 * e + q], Ka[q]);
        Kb[q] = fma(wb[e], Bb[9 * e + q], Kb[q]);
      }
    double s[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      s[i] = Ka[3 * i] * dua[0] + Ka[3 * i + 1] * dua[1] + Ka[3 * i + 2] * dua[2] + Kb[3 * i] * dub[0] +
             Kb[3 * i + 1] * dub[1] + Kb[3 * i + 2] * dub[2];
      s[i] += dpp_f64<kDppXor1>(s[i]);
      s[i] += dpp_f64<kDppXor2>(s[i]);
      s[i] += dpp_f64<kDppHalfMirror>(s[i]);
    }
    if (live && j < 3) {
      const int k = sub_load ? A.lidx[row] : -1;
      const double f = k >= 0 ? A.load_f[3 * (size_t)k + j] : 0.0;
      A.R[3 * (size_t)row + j] = (j == 0 ? s[0] : j == 1 ? s[1] : s[2]) - f;
    }
  }
}

__global__ __launch_bounds__(256) void k_box_diag_inv(BoxArgs A, const uint8_t* bcmask, double* dinv) {
  __shared__ double sB[16 * kBoxMaxTerms * 9];
  for (int q = threadIdx.x; q < 16 * kBoxMaxTerms * 9; q += 256) sB[q] = A.B[q];
  __syncthreads();
  const int j = threadIdx.x % 8;
  const int la = j, lb = j + 8;
  const int cnta = A.S->count[la], cntb = lb == kBoxDiagLane ? 0 : A.S->count[lb];
  const int da0 = A.S->d[la][0], da1 = A.S->d[la][1], da2 = A.S->d[la][2];
  const int db0 = A.S->d[lb][0], db1 = A.S->d[lb][1], db2 = A.S->d[lb][2];
  const unsigned px = A.nx + 2, py = A.ny + 2;
  const double* Ba = sB + la * (kBoxMaxTerms * 9);
  const double* Bb = sB + lb * (kBoxMaxTerms * 9);
  const int nrow_groups = (A.nrows + 31) / 32;
  for (int g = blockIdx.x; g < nrow_groups; g += gridDim.x) {
    const int row = g * 32 + (int)threadIdx.x / 8;
    const bool live = row < A.nrows;
    const int gr = live ? A.gpack[row] : 0;
    const int ri = gr & 1023, rj = (gr >> 10) & 1023, rk = gr >> 20;
    const bool nba = live && cnta > 0 && (unsigned)(ri + da0) <= (unsigned)A.nx &&
                     (unsigned)(rj + da1) <= (unsigned)A.ny && (unsigned)(rk + da2) <= (unsigned)A.nz;
    const bool nbb = live && cntb > 0 && (unsigned)(ri + db0) <= (unsigned)A.nx &&
                     (unsigned)(rj + db1) <= (unsigned)A.ny && (unsigned)(rk + db2) <= (unsigned)A.nz;
    const unsigned hr = (unsigned)(ri + 1) + px * ((unsigned)(rj + 1) + py * (unsigned)(rk + 1));
    double s[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) s[q] = 0.0;
#pragma unroll
    for (int e = 0; e < kBoxMaxTerms; ++e) {
      // absent neighbours contribute nothing (their blocks are not in the row)
      const double wa = nba ? A.W[6 * hr + (unsigned)A.S->off[la][e]] : 0.0;
      const double wb = nbb ? A.W[6 * hr + (unsigned)A.S->off[lb][e]] : 0.0;
#pragma unroll
      for (int q = 0; q < 9; ++q) s[q] = fma(wa, Ba[9 * e + q], fma(wb, Bb[9 * e + q], s[q]));
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      s[q] += dpp_f64<kDppXor1>(s[q]);
      s[q] += dpp_f64<kDppXor2>(s[q]);
      s[q] += dpp_f64<kDppHalfMirror>(s[q]);
      s[q] = -s[q];
    }
    if (live && j == 0) block_inv3_masked(s, bcmask[row], dinv + 9 * (size_t)row);
  }
}

__global__ __launch_bounds__(256) void k_boxmf_scatter(BoxMfArgs A) {
  const int px = A.nx + 2, py = A.ny + 2;
  const double beta = A.first ? 0.0 : A.scal[kBeta];
  for (int n = blockIdx.x * 256 + threadIdx.x; n < A.nrows; n += gridDim.x * 256) {
    const int g = A.gpack[n];
    const size_t hp = (size_t)((g & 1023) + 1) + (size_t)px * ((((g >> 10) & 1023) + 1) + (size_t)py * ((g >> 20) + 1));
    const uint32_t bm = A.bcmask[n];
    double v[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const size_t d = 3 * (size_t)n + i;
      const double t = A.first ? A.z[d] : fma(beta, A.p_prev[d], A.z[d]);
      v[i] = (bm >> i) & 1u ? 0.0 : t;
      if (A.p_cur) A.p_cur[d] = v[i];
    }
    A.plat[hp] = make_float4((float)v[0], (float)v[1], (float)v[2], 0.0f);
  }
}

// Lanes and stencil neighbours as in k_box_jac: lane j of a row takes neighbours j and
// j + 8 (lane 15 of the stencil is the diagonal, which the differences p_c - p_r
// absorb), so the block table reads are LDS broadcasts.  Every gather stays inside the
// padded arrays: absent terms point at offset 0 with a zero block, ghost hexes weigh
// zero, and a lane without a neighbour reads the row's own p.
__global__ __launch_bounds__(256, 6) void k_boxmf_apply(BoxMfArgs A) {
  __shared__ float sB[16 * kBoxMaxTerms * 9];
  __shared__ double sh[4];
  for (int q = threadIdx.x; q < 16 * kBoxMaxTerms * 9; q += 256) sB[q] = A.B[q];
  const int j = threadIdx.x % 8, grp = threadIdx.x / 8;
  const int la = j, lb = j + 8;
  const int cnta = A.S->count[la], cntb = lb == kBoxDiagLane ? 0 : A.S->count[lb];
  unsigned offa[kBoxMaxTerms], offb[kBoxMaxTerms];
#pragma unroll
  for (int e = 0; e < kBoxMaxTerms; ++e) {
    offa[e] = (unsigned)A.S->off[la][e];
    offb[e] = (unsigned)A.S->off[lb][e];
  }
  const int da0 = A.S->d[la][0], da1 = A.S->d[la][1], da2 = A.S->d[la][2];
  const int db0 = A.S->d[lb][0], db1 = A.S->d[lb][1], db2 = A.S->d[lb][2];
  const unsigned uoffa = (unsigned)A.S->uoff[la], uoffb = (unsigned)A.S->uoff[lb];
  const unsigned px = A.nx + 2, py = A.ny + 2;
  const float* __restrict__ W = A.W;
  const float4* __restrict__ pl = A.plat;
  __syncthreads();  // block table in LDS
  const float* Ba = sB + la * (kBoxMaxTerms * 9);
  const float* Bb = sB + lb * (kBoxMaxTerms * 9);
  const int nblk = (A.nrows + kBoxRows - 1) / kBoxRows;
  const int per = (nblk + kNumXcd - 1) / kNumXcd;
  const int x = blockIdx.x % kNumXcd, nbx = gridDim.x / kNumXcd;
  const int t_end = min((x + 1) * per, nblk);
  double dot = 0.0;
#pragma nounroll
  for (int blk = x * per + (int)blockIdx.x / kNumXcd; blk < t_end; blk += nbx) {
    // the block table is read from LDS per row, not held in registers across tiles
    asm volatile("" ::: "memory");
    const int row = blk * kBoxRows + grp;
    const bool live = row < A.nrows;
    const int gr = live ? A.gpack[row] : 0;
    const int ri = gr & 1023, rj = (gr >> 10) & 1023, rk = gr >> 20;
    const bool nba = live && cnta > 0 && (unsigned)(ri + da0) <= (unsigned)A.nx &&
                     (unsigned)(rj + da1) <= (unsigned)A.ny && (unsigned)(rk + da2) <= (unsigned)A.nz;
    const bool nbb = live && cntb > 0 && (unsigned)(ri + db0) <= (unsigned)A.nx &&
                     (unsigned)(rj + db1) <= (unsigned)A.ny && (unsigned)(rk + db2) <= (unsigned)A.nz;
    const unsigned hr = (unsigned)(ri + 1) + px * ((unsigned)(rj + 1) + py * (unsigned)(rk + 1));
    float wa[kBoxMaxTerms], wb[kBoxMaxTerms];
#pragma unroll
    for (int e = 0; e < kBoxMaxTerms; ++e) {
      wa[e] = W[6 * hr + offa[e]];
      wb[e] = W[6 * hr + offb[e]];
    }
    const float4 pr = pl[hr];
    const float4 pa = pl[nba ? hr + uoffa : hr];
    const float4 pb = pl[nbb ? hr + uoffb : hr];
    const float dua[3] = {pa.x - pr.x, pa.y - pr.y, pa.z - pr.z};
    const float dub[3] = {pb.x - pr.x, pb.y - pr.y, pb.z - pr.z};
    float Ka[9], Kb[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      Ka[q] = wa[0] * Ba[q];
      Kb[q] = wb[0] * Bb[q];
    }
#pragma unroll
    for (int e = 1; e < kBoxMaxTerms; ++e)
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        Ka[q] = fmaf(wa[e], Ba[9 * e + q], Ka[q]);
        Kb[q] = fmaf(wb[e], Bb[9 * e + q], Kb[q]);
      }
    double s[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
      s[i] = (double)(Ka[3 * i] * dua[0] + Ka[3 * i + 1] * dua[1] + Ka[3 * i + 2] * dua[2] +
                      Kb[3 * i] * dub[0] + Kb[3 * i + 1] * dub[1] + Kb[3 * i + 2] * dub[2]);
    // sum over the row's 8 lanes; every lane of the row receives the same bits
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      s[i] += dpp_f64<kDppXor1>(s[i]);
      s[i] += dpp_f64<kDppXor2>(s[i]);
      s[i] += dpp_f64<kDppHalfMirror>(s[i]);
    }
    if (live && j < 3) {
      const size_t d = 3 * (size_t)row + j;
      const double pc = A.p_cur ? A.p_cur[d] : 0.0;
      const double yv = (A.bcmask[row] >> j) & 1u ? pc : (j == 0 ? s[0] : j == 1 ? s[1] : s[2]);
      A.y[d] = yv;
      dot += pc * yv;
    }
  }
  const double t = block_reduce_all<256>(dot, sh);
  if (threadIdx.x == 0) A.partials[blockIdx.x] = t;
}

// k_boxmf_apply_kuhn (pa_box.hpp).  The row's p at the 15 lattice offsets the Kuhn tets
// reach (all components in {0, 1} or all in {0, -1}) is loaded first, the weight of tet q
// of a hex from slot q of its record; each tet's force is formed in FP32 from FP32 p and
// weights, summed in FP32 over the four tets of each path and in FP64 over the paths, in
// a fixed order.
__global__ __launch_bounds__(256) void k_boxmf_apply_kuhn(BoxMfArgs A) {
  __shared__ double sh[4];
  const int px = A.nx + 2, pxy = (A.nx + 2) * (A.ny + 2);
  const float ih[3] = {A.ih[0], A.ih[1], A.ih[2]};
  const float lam = A.lam, mu = A.mu, vol = A.vol;
  const float4* __restrict__ pl = A.plat;
  const float* __restrict__ W = A.W;
  const XcdSpan sp = xcd_span(A.nrows);
  double dot = 0.0;
  for (int row = sp.lo + sp.j * 256 + (int)threadIdx.x; row < sp.hi; row += sp.bx * 256) {
    const int g = A.gpack[row];
    const int hr = ((g & 1023) + 1) + px * ((((g >> 10) & 1023) + 1) + (A.ny + 2) * ((g >> 20) + 1));
    auto off = [&](int dx, int dy, int dz) { return dx + px * dy + pxy * dz; };
    float3 P[3][3][3];  // p at offset (dx, dy, dz) in P[dx + 1][dy + 1][dz + 1]
#pragma unroll
    for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
      for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
          const bool reach = (dx >= 0 && dy >= 0 && dz >= 0) || (dx <= 0 && dy <= 0 && dz <= 0);
          if (reach) {
            const float4 v = pl[hr + off(dx, dy, dz)];
            P[dx + 1][dy + 1][dz + 1] = make_float3(v.x, v.y, v.z);
          }
        }
    double y[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      float yq[3] = {0.0f, 0.0f, 0.0f};  // the four tets of path q, then added in FP64
#pragma unroll
      for (int m = 0; m < 4; ++m) {  // the row is local vertex m of tet q of its hex
        // offsets of the tet's four vertices relative to the row point
        int d[4][3];
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
          for (int k = 0; k < 3; ++k) d[v][k] = 0;
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
          for (int l = 0; l < 3; ++l) {
            if (l < v) d[v][kKuhnPath[q][l]] += 1;
            if (l < m) d[v][kKuhnPath[q][l]] -= 1;
          }
        const float w = W[6 * (hr + off(d[0][0], d[0][1], d[0][2])) + q];
        float G[3][3];  // G[i][j] = d p_i / d x_j
#pragma unroll
        for (int l = 0; l < 3; ++l) {
          const int c = kKuhnPath[q][l];
          const float3 a = P[d[l][0] + 1][d[l][1] + 1][d[l][2] + 1];
          const float3 b = P[d[l + 1][0] + 1][d[l + 1][1] + 1][d[l + 1][2] + 1];
          G[0][c] = (b.x - a.x) * ih[c];
          G[1][c] = (b.y - a.y) * ih[c];
          G[2][c] = (b.z - a.z) * ih[c];
        }
        const float tr = G[0][0] + G[1][1] + G[2][2];
        float S[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i) S[i][i] = fmaf(2.0f * mu, G[i][i], lam * tr);
        S[0][1] = S[1][0] = mu * (G[0][1] + G[1][0]);
        S[0][2] = S[2][0] = mu * (G[0][2] + G[2][0]);
        S[1][2] = S[2][1] = mu * (G[1][2] + G[2][1]);
        float f[3] = {0.0f, 0.0f, 0.0f};
        if (m > 0) {
          const int c = kKuhnPath[q][m - 1];
#pragma unroll
          for (int i = 0; i < 3; ++i) f[i] = S[i][c] * ih[c];
        }
        if (m < 3) {
          const int c = kKuhnPath[q][m];
#pragma unroll
          for (int i = 0; i < 3; ++i) f[i] = fmaf(-S[i][c], ih[c], f[i]);
        }
        const float vw = vol * w;
#pragma unroll
        for (int i = 0; i < 3; ++i) yq[i] = fmaf(vw, f[i], yq[i]);
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) y[i] += (double)yq[i];
    }
    const uint32_t bm = A.bcmask[row];
    // constrained rows: y = p there, which is 0 (masked); the p.y partial from the FP32 p
    // the operator acts on (its rounding moves p.y by ~1e-7 relative, below the operator's
    // own FP32 error), so the FP64 p is not read
    const float3 pr = P[1][1][1];
    const double pf[3] = {pr.x, pr.y, pr.z};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const size_t dd = 3 * (size_t)row + i;
      const double yv = (bm >> i) & 1u ? 0.0 : y[i];
      A.y[dd] = yv;
      dot += pf[i] * yv;
    }
  }
  const double t = block_reduce_all<256>(dot, sh);
  if (threadIdx.x == 0) A.partials[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void k_box_diag_inv_kuhn(BoxArgs A, const uint8_t* bcmask, double* dinv) {
  const int px = A.nx + 2, pxy = (A.nx + 2) * (A.ny + 2);
  const double ih[3] = {A.ih[0], A.ih[1], A.ih[2]};
  const double lm = A.m.c01 + A.m.c33, mu = A.m.c33, vol = A.vol;
  int tp[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) tp[q] = A.tperm[q];
  const XcdSpan sp = xcd_span(A.nrows);
  for (int row = sp.lo + sp.j * 256 + (int)threadIdx.x; row < sp.hi; row += sp.bx * 256) {
    const int gp = A.gpack[row];
    const int hr = ((gp & 1023) + 1) + px * ((((gp >> 10) & 1023) + 1) + (A.ny + 2) * ((gp >> 20) + 1));
    double s[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) s[k] = 0.0;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
#pragma unroll
      for (int m = 0; m < 4; ++m) {  // the row is local vertex m of tet q of its hex
        int d0[3] = {0, 0, 0};       // the hex's lower corner relative to the row point
#pragma unroll
        for (int l = 0; l < 3; ++l)
          if (l < m) d0[kKuhnPath[q][l]] -= 1;
        const double w = A.W[6 * (hr + d0[0] + px * d0[1] + pxy * d0[2]) + tp[q]];
        double g[3] = {0.0, 0.0, 0.0};
        if (m > 0) g[kKuhnPath[q][m - 1]] += ih[kKuhnPath[q][m - 1]];
        if (m < 3) g[kKuhnPath[q][m]] -= ih[kKuhnPath[q][m]];
        const double vw = vol * w, gg = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int b = 0; b < 3; ++b) s[3 * a + b] = fma(vw, lm * g[a] * g[b] + (a == b ? mu * gg : 0.0), s[3 * a + b]);
      }
    }
    block_inv3_masked(s, bcmask[row], dinv + 9 * (size_t)row);
  }
}

__global__ __launch_bounds__(256) void k_box_res_kuhn(BoxArgs A, int sub_load) {
  const int px = A.nx + 2, pxy = (A.nx + 2) * (A.ny + 2);
  const double ih[3] = {A.ih[0], A.ih[1], A.ih[2]};
  const double lam = A.m.c01, mu = A.m.c33, vol = A.vol;
  int tp[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) tp[q] = A.tperm[q];
  const double* __restrict__ ul = A.ulat;
  const double* __restrict__ W = A.W;
  const XcdSpan sp = xcd_span(A.nrows);
  for (int row = sp.lo + sp.j * 256 + (int)threadIdx.x; row < sp.hi; row += sp.bx * 256) {
    const int g = A.gpack[row];
    const int hr = ((g & 1023) + 1) + px * ((((g >> 10) & 1023) + 1) + (A.ny + 2) * ((g >> 20) + 1));
    auto off = [&](int dx, int dy, int dz) { return dx + px * dy + pxy * dz; };
    double3 P[3][3][3];  // u at offset (dx, dy, dz) in P[dx + 1][dy + 1][dz + 1]
#pragma unroll
    for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
      for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
          const bool reach = (dx >= 0 && dy >= 0 && dz >= 0) || (dx <= 0 && dy <= 0 && dz <= 0);
          if (reach) {
            const double* v = ul + 3 * (size_t)(hr + off(dx, dy, dz));
            P[dx + 1][dy + 1][dz + 1] = make_double3(v[0], v[1], v[2]);
          }
        }
    double y[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 6; ++q) {
#pragma unroll
      for (int m = 0; m < 4; ++m) {  // the row is local vertex m of tet q of its hex
        int d[4][3];
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
          for (int k = 0; k < 3; ++k) d[v][k] = 0;
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
          for (int l = 0; l < 3; ++l) {
            if (l < v) d[v][kKuhnPath[q][l]] += 1;
            if (l < m) d[v][kKuhnPath[q][l]] -= 1;
          }
        const double w = W[6 * (hr + off(d[0][0], d[0][1], d[0][2])) + tp[q]];
        double G[3][3];  // G[i][j] = d u_i / d x_j
#pragma unroll
        for (int l = 0; l < 3; ++l) {
          const int c = kKuhnPath[q][l];
          const double3 a = P[d[l][0] + 1][d[l][1] + 1][d[l][2] + 1];
          const double3 b = P[d[l + 1][0] + 1][d[l + 1][1] + 1][d[l + 1][2] + 1];
          G[0][c] = (b.x - a.x) * ih[c];
          G[1][c] = (b.y - a.y) * ih[c];
          G[2][c] = (b.z - a.z) * ih[c];
        }
        const double tr = G[0][0] + G[1][1] + G[2][2];
        double S[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i) S[i][i] = fma(2.0 * mu, G[i][i], lam * tr);
        S[0][1] = S[1][0] = mu * (G[0][1] + G[1][0]);
        S[0][2] = S[2][0] = mu * (G[0][2] + G[2][0]);
        S[1][2] = S[2][1] = mu * (G[1][2] + G[2][1]);
        double f[3] = {0.0, 0.0, 0.0};
        if (m > 0) {
          const int c = kKuhnPath[q][m - 1];
#pragma unroll
          for (int i = 0; i < 3; ++i) f[i] = S[i][c] * ih[c];
        }
        if (m < 3) {
          const int c = kKuhnPath[q][m];
#pragma unroll
          for (int i = 0; i < 3; ++i) f[i] = fma(-S[i][c], ih[c], f[i]);
        }
        const double vw = vol * w;
#pragma unroll
        for (int i = 0; i < 3; ++i) y[i] = fma(vw, f[i], y[i]);
      }
    }
    const int k = sub_load ? A.lidx[row] : -1;
#pragma unroll
    for (int i = 0; i < 3; ++i)
      A.R[3 * (size_t)row + i] = y[i] - (k >= 0 ? A.load_f[3 * (size_t)k + i] : 0.0);
  }
}

}  // namespace pa
