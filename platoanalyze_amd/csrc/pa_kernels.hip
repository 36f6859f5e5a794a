// Chunked element kernels: one work-group owns a run of consecutive block rows,
// accumulates the contributions of every element touching those rows in LDS
// (atomic-free in HBM: each CRS value and residual entry is written exactly once,
// coalesced), then streams the finished segment out.
#include "pa_kernels.hpp"
#include "pa_fixed.hpp"
#include "pa_dpp.hpp"

namespace pa {

constexpr int kNoSlotDev = 0xFFFF;  // no transposed partner in the chunk

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

// Scale of a chunk's fixed-point Jacobian accumulation (lds_add_fixed64): the largest power of two S with
// S * bound < 2^57, bound = gbound (|lambda| + 2 |mu|) on every element block entry
// (weights <= 1), so up to 64 such contributions per entry cannot overflow and each is
// resolved to 2^-57 of the bound.
__device__ __forceinline__ double fixed_scale(const ChunkArgs& A, int c) {
  const double B = (double)A.gbound[c] * (fabs(A.m.c01) + 2.0 * fabs(A.m.c33));
  return B > 0.0 ? ldexp(1.0, 56 - ilogb(B)) : 1.0;
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

// select x[a] for a runtime a in 0..3 without dynamic register indexing
__device__ __forceinline__ double sel4(int a, double x0, double x1, double x2, double x3) {
  return a == 0 ? x0 : (a == 1 ? x1 : (a == 2 ? x2 : x3));
}
__device__ __forceinline__ int sel4i(int a, int x0, int x1, int x2, int x3) {
  return a == 0 ? x0 : (a == 1 ? x1 : (a == 2 ? x2 : x3));
}


// Vector-valued element kernels (residual, dR/dz, energy and its gradient) over the
// same chunks: node coordinates, density and state staged in LDS, elements streamed
// as records of local node indices, owned rows accumulated in LDS.
template <int BITS, int MODE, int NT>
__global__ __launch_bounds__(NT) void k_chunk(ChunkArgs A) {
  extern __shared__ double lds[];
  constexpr bool kRes = MODE == kModeResidual;
  constexpr bool kGzT = MODE == kModeGradZT;
  constexpr bool kGradZ = MODE == kModeEnergyGradZ;
  constexpr bool kPnV = MODE == kModePNorm, kPnG = MODE == kModePNormGrad, kPnX = MODE == kModePNormGradX;
  constexpr bool kAdj = MODE == kModeAdjZ;
  constexpr bool kGx = MODE == kModeGradX, kEGx = MODE == kModeEnergyGradX, kVol = MODE == kModeVolume;
  constexpr bool kReduce = MODE == kModeEnergy || kGradZ || kPnV || kVol;
  constexpr int kBlkW = kGzT ? 3 : (kGx ? 9 : 0);                      // doubles per block in LDS
  constexpr int kRowW = (kRes || kPnG || kEGx || kVol || kPnX) ? 3 : ((kGradZ || kAdj) ? 1 : 0);  // doubles per row
  // second per-row output (the residual: the low words of its two-word fixed-point sums)
  constexpr int kRow2 = kRes ? 3 : ((kPnG || kVol) ? 1 : 0);

  const int c = xcd_chunk(blockIdx.x, A.nchunks);
  const int r0 = A.chunk_rows[c], r1 = A.chunk_rows[c + 1];
  const int nr = r1 - r0;
  const int b0 = A.rowmap[r0];
  const int nb = A.rowmap[r1] - b0;
  const int n0 = A.node_offs[c], nnode = A.node_offs[c + 1] - n0;
  double4* nodes = reinterpret_cast<double4*>(lds);
  double* ul = reinterpret_cast<double*>(nodes + A.max_nodes);
  double* ul2 = ul + 3 * (size_t)A.max_nodes;
  double* seg = ul2 + (kAdj ? 3 * (size_t)A.max_nodes : 0);
  double* res = seg + (size_t)A.max_blocks * kBlkW;
  double* res2 = res + (size_t)A.max_rows * kRowW;
  int* roff = reinterpret_cast<int*>(res2 + (size_t)A.max_rows * kRow2);

  // the residual sums in fixed point (order-independent): the scale comes from the
  // largest |u| of the chunk's nodes, a NaN among them poisons it
  __shared__ double red[NT / 64];
  double umax = 0.0;
  for (int k = threadIdx.x; k < nnode; k += NT) {
    const size_t gid = (size_t)A.nodes[n0 + k];
    nodes[k] = make_double4(A.X[3 * gid], A.X[3 * gid + 1], A.X[3 * gid + 2], A.z[gid]);
    ul[3 * k + 0] = A.u[3 * gid + 0];
    ul[3 * k + 1] = A.u[3 * gid + 1];
    ul[3 * k + 2] = A.u[3 * gid + 2];
    if constexpr (kRes) {
#pragma unroll
      for (int i = 0; i < 3; ++i) umax = max_nan(umax, fabs(ul[3 * k + i]));
    }
    if constexpr (kAdj) {
      ul2[3 * k + 0] = A.u2[3 * gid + 0];
      ul2[3 * k + 1] = A.u2[3 * gid + 1];
      ul2[3 * k + 2] = A.u2[3 * gid + 2];
    }
  }
  if constexpr (kBlkW > 0) {
    for (int k = threadIdx.x; k < nb * kBlkW; k += NT) seg[k] = 0.0;
    for (int k = threadIdx.x; k <= nr; k += NT) roff[k] = A.rowmap[r0 + k] - b0;
  }
  if constexpr (kRowW > 0)
    for (int k = threadIdx.x; k < nr * kRowW; k += NT) res[k] = 0.0;
  if constexpr (kRow2 > 0)
    for (int k = threadIdx.x; k < nr * kRow2; k += NT) res2[k] = 0.0;
  if constexpr (kRes) block_max_write<NT>(umax, red);
  __syncthreads();
  // |contribution| <= vol |grad N|^2 * 3 * (|c00| + 2|c01| + |c33|) * 8 max|u| for weights <= 1
  double S = 1.0;
  if constexpr (kRes) {
    const double cs = fabs(A.m.c00) + 2.0 * fabs(A.m.c01) + fabs(A.m.c33);
    S = fixed_scale_54(24.0 * cs * (double)A.gbound[c] * block_max_read<NT>(red));
  }

  double energy = 0.0;
  const int l0 = A.list_offs[c], l1 = A.list_offs[c + 1];
  for (int k = l0 + threadIdx.x; k < l1; k += NT) {
    const uint2 lw = reinterpret_cast<const uint2*>(A.ent_ln)[k];
    const int ln[4] = {(int)(lw.x & 0xFFFFu), (int)(lw.x >> 16), (int)(lw.y & 0xFFFFu), (int)(lw.y >> 16)};
    double x[4][3], g[4][3], u[4][3], vol, rho = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const double4 p = nodes[ln[a]];
      x[a][0] = p.x;
      x[a][1] = p.y;
      x[a][2] = p.z;
      rho += p.w;
#pragma unroll
      for (int i = 0; i < 3; ++i) u[a][i] = ul[3 * ln[a] + i];
    }
    tet_gradients(x, g, vol);
    rho = rho / 4.0;
    const double w = penalty_w(A.m, rho);
    double eps[6], s[6];
    tet_strain(u, g, eps);
    iso_stress(A.m, eps, s);

    if constexpr (MODE == kModeEnergy || kGradZ) {
      // ScalarProduct<6>(sigma_w, eps) * 0.5 * vol; counted by the chunk owning vertex 0
      double se = 0.0;
#pragma unroll
      for (int q = 0; q < 6; ++q) se += (w * s[q]) * eps[q];
      if (ln[0] < nr) energy += 0.5 * se * vol;
      if constexpr (kGradZ) {
        // d(0.5 vol sigma_w:eps)/dz_a = 0.5 vol (sigma:eps) * dw/drho * 1/4
        double sr = 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q) sr += s[q] * eps[q];
        const double dw = penalty_dw(A.m, rho);
        const double gz = 0.5 * vol * sr * dw * 0.25;
#pragma unroll
        for (int a = 0; a < 4; ++a)
          if (ln[a] < nr) lds_add(&res[ln[a]], gz);
      }
    }
    if constexpr (kPnV || kPnG || kPnX) {
      // measure m and, for the gradient, the stress-space direction t with
      // dm/du = c_u * B^T C0 t and dm/drho = c_z * m
      double mval, t[6], cu, cz;
      if (A.measure == 0) {
        double ss = 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q) ss += (w * s[q]) * (w * s[q]);
        mval = sqrt(ss);
#pragma unroll
        for (int q = 0; q < 6; ++q) t[q] = w * s[q];  // d|sigma_w|/du = w B^T C0 sigma_w / m
        cu = mval > 0.0 ? w / mval : 0.0;
        cz = penalty_dw(A.m, rho) / w;
      } else {
        const double d01 = s[0] - s[1], d12 = s[1] - s[2], d20 = s[2] - s[0];
        const double j2 = 0.5 * (d01 * d01 + d12 * d12 + d20 * d20) + 3.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
        const double vm0 = sqrt(j2), rq = pow(rho, A.q);
        mval = rq * vm0;
        t[0] = 2.0 * s[0] - s[1] - s[2];
        t[1] = 2.0 * s[1] - s[2] - s[0];
        t[2] = 2.0 * s[2] - s[0] - s[1];
        t[3] = 6.0 * s[3];
        t[4] = 6.0 * s[4];
        t[5] = 6.0 * s[5];
        cu = vm0 > 0.0 ? rq / (2.0 * vm0) : 0.0;  // d vm / d sigma = dJ/dsigma / (2 vm)
        cz = A.q / rho;
      }
      const double mp = pow(mval, A.pn);
      if constexpr (kPnV) {
        if (ln[0] < nr) energy += mp * vol;
      } else if constexpr (kPnX) {
        // dS/dX_{b,k} = vol m^p g_bk (dvol/dX_{b,k} = vol g_bk)
        //             + vol p m^(p-1) c_u (C0 t) . deps_{b,k},
        // deps_{b,k}: Voigt strain (engineering shear) of dH/dX_{b,k} = -H[:,k] g_b^T,
        // H = sum_a u_a g_a^T; rho does not depend on X.
        double H[3][3];
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
          for (int q2 = 0; q2 < 3; ++q2)
            H[p][q2] = u[0][p] * g[0][q2] + u[1][p] * g[1][q2] + u[2][p] * g[2][q2] + u[3][p] * g[3][q2];
        double tau[6];
        iso_stress(A.m, t, tau);
        const double mpm1 = mval > 0.0 ? mp / mval : 0.0;
        const double fv = A.pscale * vol * mp;
        const double fs = A.pscale * vol * A.pn * mpm1 * cu;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          if (ln[b] >= nr) continue;
          double* dst = &res[3 * ln[b]];
#pragma unroll
          for (int k2 = 0; k2 < 3; ++k2) {
            const double de = -(tau[0] * H[0][k2] * g[b][0] + tau[1] * H[1][k2] * g[b][1] + tau[2] * H[2][k2] * g[b][2] +
                                tau[3] * (H[1][k2] * g[b][2] + H[2][k2] * g[b][1]) +
                                tau[4] * (H[0][k2] * g[b][2] + H[2][k2] * g[b][0]) +
                                tau[5] * (H[0][k2] * g[b][1] + H[1][k2] * g[b][0]));
            lds_add(dst + k2, fv * g[b][k2] + fs * de);
          }
        }
      } else {
        // dS/du_a = vol p m^(p-1) c_u B_a^T C0 t ; dS/dz_a = vol p m^p c_z / 4
        const double mpm1 = mval > 0.0 ? mp / mval : 0.0;
        const double fu = A.pscale * vol * A.pn * mpm1 * cu;
        const double fz = A.pscale * vol * A.pn * mp * cz * 0.25;
        double tau[6];
        iso_stress(A.m, t, tau);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          if (ln[a] < nr) {
            double r[3];
            tet_divergence(tau, g[a], fu, r);
            double* dst = &res[3 * ln[a]];
            lds_add(dst + 0, r[0]);
            lds_add(dst + 1, r[1]);
            lds_add(dst + 2, r[2]);
            lds_add(&res2[ln[a]], fz);
          }
        }
      }
    }
    if constexpr (kAdj) {
      // (dR/dz)^T lambda: dw/drho/4 * vol * eps(lambda_e) . sigma0(u_e)
      double l[4][3], el[6];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int i = 0; i < 3; ++i) l[a][i] = ul2[3 * ln[a] + i];
      tet_strain(l, g, el);
      double dot = 0.0;
#pragma unroll
      for (int q = 0; q < 6; ++q) dot += el[q] * s[q];
      const double dw = penalty_dw(A.m, rho);
      const double v = 0.25 * dw * vol * dot;
#pragma unroll
      for (int a = 0; a < 4; ++a)
        if (ln[a] < nr) lds_add(&res[ln[a]], v);
    }
    if constexpr (kRes) {
      double sw[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) sw[q] = w * s[q];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        if (ln[a] < nr) {
          double r[3];
          tet_divergence(sw, g[a], vol, r);
          double* dst = &res[3 * ln[a]];
          double* lo = &res2[3 * ln[a]];
          lds_add_fixed128(dst + 0, lo + 0, r[0], S);
          lds_add_fixed128(dst + 1, lo + 1, r[1], S);
          lds_add_fixed128(dst + 2, lo + 2, r[2], S);
        }
      }
    }
    if constexpr (kGzT) {
      // dR_b/dz_a = dw/drho * 1/4 * (vol B_b^T sigma); rows are the control nodes a
      uint32_t cw[4];
      if constexpr (BITS == 4) {
        const uint2 t = reinterpret_cast<const uint2*>(A.ent_code)[k];
        cw[0] = t.x;
        cw[1] = t.y;
        cw[2] = t.y;
        cw[3] = t.y;
      } else {
        const uint4 t = reinterpret_cast<const uint4*>(A.ent_code)[k];
        cw[0] = t.x;
        cw[1] = t.y;
        cw[2] = t.z;
        cw[3] = t.w;
      }
      const double dw = penalty_dw(A.m, rho) * 0.25;
      double rb[4][3];
#pragma unroll
      for (int b = 0; b < 4; ++b) tet_divergence(s, g[b], vol, rb[b]);
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        if (ln[a] >= nr) continue;
        const uint32_t code = BITS == 4 ? (cw[a >> 1] >> (16 * (a & 1))) : cw[a];
        const int base = roff[ln[a]];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int sl = base + (int)((code >> (BITS * b)) & ((1u << BITS) - 1));
#pragma unroll
          for (int i = 0; i < 3; ++i) lds_add(&seg[3 * sl + i], dw * rb[b][i]);
        }
      }
    }
    if constexpr (kVol) {
      // Geometric volume V = sum_e vol_e w(rho_e) with the criterion's own penalty (in
      // A.m): dV/dz_a = vol w'(rho) / 4, dV/dX_a = w vol g_a (dvol/dX_a = vol g_a)
      const double dw = penalty_dw(A.m, rho);
      if (ln[0] < nr) energy += vol * w;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        if (ln[a] >= nr) continue;
        double* dst = &res[3 * ln[a]];
#pragma unroll
        for (int k2 = 0; k2 < 3; ++k2) lds_add(dst + k2, w * vol * g[a][k2]);
        lds_add(&res2[ln[a]], 0.25 * vol * dw);
      }
    }
    if constexpr (kGx || kEGx) {
      // Shape derivatives.  With H = sum_b u_b g_b^T, sigma_w = w C eps (3x3 tensor S),
      // dvol/dX_{b,k} = vol g_bk, dg_aj/dX_{b,k} = -g_bj g_ak, dH/dX_{b,k} = -H[:,k] g_b^T:
      //   energy: dE/dX_{a,k} = vol (psi g_ak - sum_p H_pk (S g_a)_p), psi = sigma_w.eps / 2
      //   residual r_a = vol S g_a:
      //   dr_{a,i}/dX_{b,k} = vol (g_bk (S g_a)_i - (S g_b)_i g_ak + (dS_bk g_a)_i),
      //   dS_bk = w C sym(dH/dX_{b,k}).
      double H[3][3];
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < 3; ++q) H[p][q] = u[0][p] * g[0][q] + u[1][p] * g[1][q] + u[2][p] * g[2][q] + u[3][p] * g[3][q];
      const double S[3][3] = {{w * s[0], w * s[5], w * s[4]}, {w * s[5], w * s[1], w * s[3]}, {w * s[4], w * s[3], w * s[2]}};
      double Sg[4][3];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int i = 0; i < 3; ++i) Sg[a][i] = S[i][0] * g[a][0] + S[i][1] * g[a][1] + S[i][2] * g[a][2];
      if constexpr (kEGx) {
        double psi = 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q) psi += (w * s[q]) * eps[q];
        psi *= 0.5;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          if (ln[a] >= nr) continue;
          double* dst = &res[3 * ln[a]];
#pragma unroll
          for (int k2 = 0; k2 < 3; ++k2)
            lds_add(dst + k2, vol * (psi * g[a][k2] - (H[0][k2] * Sg[a][0] + H[1][k2] * Sg[a][1] + H[2][k2] * Sg[a][2])));
        }
      } else {
        uint32_t cw[4];
        if constexpr (BITS == 4) {
          const uint2 t = reinterpret_cast<const uint2*>(A.ent_code)[k];
          cw[0] = t.x;
          cw[1] = t.y;
          cw[2] = t.y;
          cw[3] = t.y;
        } else {
          const uint4 t = reinterpret_cast<const uint4*>(A.ent_code)[k];
          cw[0] = t.x;
          cw[1] = t.y;
          cw[2] = t.z;
          cw[3] = t.w;
        }
#pragma unroll
        for (int b = 0; b < 4; ++b) {  // config vertex b = the chunk's row
          if (ln[b] >= nr) continue;
          const uint32_t code = BITS == 4 ? (cw[b >> 1] >> (16 * (b & 1))) : cw[b];
          const int base = roff[ln[b]];
#pragma unroll
          for (int k2 = 0; k2 < 3; ++k2) {
            // dH = -H[:,k] g_b^T; Voigt strain with engineering shear; dS = w C deps
            double de[6];
            de[0] = -H[0][k2] * g[b][0];
            de[1] = -H[1][k2] * g[b][1];
            de[2] = -H[2][k2] * g[b][2];
            de[3] = -(H[1][k2] * g[b][2] + H[2][k2] * g[b][1]);
            de[4] = -(H[0][k2] * g[b][2] + H[2][k2] * g[b][0]);
            de[5] = -(H[0][k2] * g[b][1] + H[1][k2] * g[b][0]);
            double ds[6];
            iso_stress(A.m, de, ds);
            const double dS[3][3] = {{w * ds[0], w * ds[5], w * ds[4]}, {w * ds[5], w * ds[1], w * ds[3]},
                                     {w * ds[4], w * ds[3], w * ds[2]}};
#pragma unroll
            for (int a = 0; a < 4; ++a) {
              const int sl = base + (int)((code >> (BITS * a)) & ((1u << BITS) - 1));
#pragma unroll
              for (int i = 0; i < 3; ++i)
                lds_add(&seg[9 * sl + 3 * k2 + i],
                        vol * (g[b][k2] * Sg[a][i] - Sg[b][i] * g[a][k2] +
                               (dS[i][0] * g[a][0] + dS[i][1] * g[a][1] + dS[i][2] * g[a][2])));
            }
          }
        }
      }
    }
  }

  if constexpr (kReduce) {
    // block reduction of the energy / p-norm sum (fixed order -> deterministic)
    __shared__ double wsum[NT / 64];
    double v = energy;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int q = 0; q < NT / 64; ++q) t += wsum[q];
      A.partials[c] = t;
    }
  }
  __syncthreads();
  if constexpr (kRes) {
    const double inv = 1.0 / S;  // a power of two: exact (NaN for a non-finite u)
    for (int k = threadIdx.x; k < nr * 3; k += NT) res[k] = S == S ? fixed_value2(res[k], res2[k], inv) : NAN;
    __syncthreads();
    if (A.sub_load) {
      for (int k = A.load_offs[c] + threadIdx.x; k < A.load_offs[c + 1]; k += NT) {
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
    for (int k = threadIdx.x; k < nb * kBlkW; k += NT) dst[k] = seg[k];
  }
  if constexpr (kRowW > 0) {
    double* dst = A.out_vec + (size_t)r0 * kRowW;
    for (int k = threadIdx.x; k < nr * kRowW; k += NT) dst[k] = res[k];
  }
  if constexpr (kPnG || kVol) {
    double* dst = A.out_vec2 + (size_t)r0;
    for (int k = threadIdx.x; k < nr; k += NT) dst[k] = res2[k];
  }
}

// Local edges (a,b), a<b, of a tet.
__device__ constexpr int kEdgeA[6] = {0, 0, 0, 1, 1, 2};
__device__ constexpr int kEdgeB[6] = {1, 2, 3, 2, 3, 3};

__device__ __forceinline__ int max16(int v) {
  v = max(v, __builtin_amdgcn_update_dpp(0, v, kDppXor1, 0xF, 0xF, false));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, kDppXor2, 0xF, 0xF, false));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, kDppHalfMirror, 0xF, 0xF, false));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, kDppMirror, 0xF, 0xF, false));
  return v;
}

// Slot code of block (row vertex x, column vertex y) of an element entry.
template <int BITS>
__device__ __forceinline__ uint32_t slot_code(const uint32_t cw[4], int x, int y) {
  if constexpr (BITS == 4) return (cw[x >> 1] >> (16 * (x & 1) + 4 * y)) & 15u;
  else return (cw[x] >> (8 * y)) & 255u;
}

// LDS layout of the Jacobian kernel (byte offsets, 16-byte aligned sections).
struct JacLds {
  size_t seg, nodes, roff, col, fl, total;
};
__host__ __device__ inline JacLds jac_lds_layout(int max_rows, int max_blocks, int max_nodes) {
  auto al = [](size_t x) { return (x + 15) & ~(size_t)15; };
  JacLds L;
  L.seg = 0;
  L.nodes = al(L.seg + sizeof(double) * 9 * (size_t)max_blocks);
  L.roff = al(L.nodes + sizeof(double) * 4 * (size_t)max_nodes);
  L.col = al(L.roff + sizeof(int) * ((size_t)max_rows + 1));
  L.fl = al(L.col + sizeof(int) * (size_t)max_blocks);
  L.total = al(L.fl + sizeof(double) * 3 * (size_t)max_rows);
  return L;
}

// Jacobian (+ residual) of one chunk of rows, accumulated per edge.
//   * The chunk's node table (its rows, then the nodes they couple to) is staged in
//     LDS once with near-coalesced loads; elements arrive as streamed records of
//     four local node indices plus slot codes, so the element loop does no global
//     gathers.  A vertex is owned by the chunk iff its local index < number of rows.
//   * Off-diagonal blocks are accumulated once per element edge: in the row of the
//     smaller vertex number when both ends are owned (the other block is its
//     transpose, K being symmetric), else in the row of the owned end.  The six
//     edges are unrolled, so vertex pairs are compile-time constants.
//   * Diagonal blocks use the rigid-translation identity of every element,
//     sum_b K_ab = 0 (sum_b grad N_b = 0): K_rr = -sum_{c != r} K_rc.
//   * Residual R_r = sum_c K_rc u_c - f_r = sum_{c != r} K_rc (u_c - u_r) - f_r.
// Transposes (slot of the partner block from a setup table), diagonals and residual
// are finished by 16-lane row groups with DPP reductions, then the chunk's CRS
// segment is streamed out once.
// Rows, blocks and node table of one chunk.
struct JacChunk {
  int c, r0, nr, b0, nb, n0, nnode;
};
__device__ __forceinline__ JacChunk jac_chunk_info(const ChunkArgs& A, int c) {
  JacChunk J;
  J.c = c;
  J.r0 = A.chunk_rows[c];
  J.nr = A.chunk_rows[c + 1] - J.r0;
  J.b0 = A.rowmap[J.r0];
  J.nb = A.rowmap[J.r0 + J.nr] - J.b0;
  J.n0 = A.node_offs[c];
  J.nnode = A.node_offs[c + 1] - J.n0;
  return J;
}

// Phase 1 (before a barrier): zero the chunk's CRS segment, stage row offsets, columns
// and the node table; zero the load rows.
template <int NT, bool WITH_R>
__device__ __forceinline__ void jac_stage(const ChunkArgs& A, const JacChunk& J, double* seg, double4* nodes,
                                          int* roff, int* col, double* fl) {
  {
    // seg starts 16-byte aligned and holds 9*nb doubles
    double2* s2 = reinterpret_cast<double2*>(seg);
    const int n2 = (9 * J.nb + 1) / 2;
    for (int k = threadIdx.x; k < n2; k += NT) s2[k] = make_double2(0.0, 0.0);
  }
  for (int k = threadIdx.x; k <= J.nr; k += NT) roff[k] = A.rowmap[J.r0 + k] - J.b0;
  for (int k = threadIdx.x; k < J.nb; k += NT) col[k] = A.colidx[J.b0 + k];
  for (int k = threadIdx.x; k < J.nnode; k += NT) {
    const size_t gid = (size_t)A.nodes[J.n0 + k];
    nodes[k] = make_double4(A.X[3 * gid], A.X[3 * gid + 1], A.X[3 * gid + 2], A.z[gid]);
  }
  if constexpr (WITH_R)
    for (int k = threadIdx.x; k < J.nr * 3; k += NT) fl[k] = 0.0;
}

// Phase 2 (after the staging barrier): the chunk's traction loads, then the elements.
template <int BITS, int NT, bool WITH_R>
__device__ __forceinline__ void jac_elements(const ChunkArgs& A, const JacChunk& J, double* seg,
                                             const double4* nodes, const int* roff, double* fl) {
  const int c = J.c, nr = J.nr;
  if constexpr (WITH_R) {
    if (A.sub_load)
      for (int k = A.load_offs[c] + threadIdx.x; k < A.load_offs[c + 1]; k += NT) {
        const int lr = A.load_row[k];
        fl[3 * lr + 0] = A.load_f[3 * k + 0];
        fl[3 * lr + 1] = A.load_f[3 * k + 1];
        fl[3 * lr + 2] = A.load_f[3 * k + 2];
      }
  }
  const int l0 = A.list_offs[c], l1 = A.list_offs[c + 1];
  const double S = fixed_scale(A, c);
  for (int k = l0 + threadIdx.x; k < l1; k += NT) {
    const uint2 lw = reinterpret_cast<const uint2*>(A.ent_ln)[k];
    const int ln[4] = {(int)(lw.x & 0xFFFFu), (int)(lw.x >> 16), (int)(lw.y & 0xFFFFu), (int)(lw.y >> 16)};
    uint32_t cw[4];
    if constexpr (BITS == 4) {
      const uint2 t = reinterpret_cast<const uint2*>(A.ent_code)[k];
      cw[0] = t.x;
      cw[1] = t.y;
      cw[2] = t.y;
      cw[3] = t.y;
    } else {
      const uint4 t = reinterpret_cast<const uint4*>(A.ent_code)[k];
      cw[0] = t.x;
      cw[1] = t.y;
      cw[2] = t.z;
      cw[3] = t.w;
    }
    double x[4][3], g[4][3], vol, rho = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const double4 p = nodes[ln[a]];
      x[a][0] = p.x;
      x[a][1] = p.y;
      x[a][2] = p.z;
      rho += p.w;
    }
    tet_gradients(x, g, vol);
    rho = rho / 4.0;
    const double w = penalty_w(A.m, rho);
    const double vl = vol * w * A.m.c01, vm = vol * w * A.m.c33;
    bool own[4];
    int ro[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      own[a] = ln[a] < nr;
      ro[a] = own[a] ? roff[ln[a]] : 0;
    }
#pragma unroll
    for (int t = 0; t < 6; ++t) {
      const int a = kEdgeA[t], b = kEdgeB[t];
      if (!(own[a] || own[b])) continue;
      // row b when a is not owned, or both are and b has the smaller vertex number
      // (local order = global order for rows)
      const bool rowb = !own[a] || (own[b] && ln[b] < ln[a]);
      const int slot = rowb ? ro[b] + (int)slot_code<BITS>(cw, b, a) : ro[a] + (int)slot_code<BITS>(cw, a, b);
      double* dst = &seg[9 * slot];
      // K_ab = vw (lam g_a g_b^T + mu g_b g_a^T + mu (g_a.g_b) I); row b stores K_ab^T
      const double d = vm * (g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2]);
      double kv[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) kv[i][j] = (vl * g[a][i]) * g[b][j] + (vm * g[b][i]) * g[a][j] + (i == j ? d : 0.0);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) lds_add_fixed64(dst + 3 * i + j, rowb ? kv[j][i] : kv[i][j], S);
    }
  }
}

// Phase 2b (after the element barrier): the fixed-point segment back to doubles.
template <int NT>
__device__ __forceinline__ void jac_fixed_to_double(double* seg, int n, double S) {
  const double inv = 1.0 / S;  // a power of two: exact
  for (int k = threadIdx.x; k < n; k += NT) seg[k] = fixed_value(seg[k], inv);
}

// Phase 3 (after the element barrier): row groups finish the rows.
template <int NT, bool WITH_R>
__device__ __forceinline__ void jac_rows(const ChunkArgs& A, const JacChunk& J, double* seg, const int* roff,
                                         const int* col, const double* fl) {
  const int r0 = J.r0, nr = J.nr, b0 = J.b0;
  // Row groups of 16 lanes finish one row each: transposed blocks, diagonal, residual.
  constexpr int kG = 16;
  const int lane = threadIdx.x % kG;
  const uint16_t* tsl = A.tslot + b0;
  for (int lr = threadIdx.x / kG; lr < nr; lr += NT / kG) {
    const int rr = r0 + lr;
    double sum[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) sum[q] = 0.0;
    double ru[3] = {0.0, 0.0, 0.0}, ur[3] = {0.0, 0.0, 0.0};
    if constexpr (WITH_R) {
      ur[0] = A.u[3 * (size_t)rr];
      ur[1] = A.u[3 * (size_t)rr + 1];
      ur[2] = A.u[3 * (size_t)rr + 2];
    }
    int dslot = -1;
    for (int s2 = roff[lr] + lane; s2 < roff[lr + 1]; s2 += kG) {
      const int cc = col[s2];
      if (cc == rr) {
        dslot = s2;
        continue;
      }
      double blk[9];
      const int p = tsl[s2];
      if (p != kNoSlotDev) {  // both rows in the chunk, below the diagonal: transpose of (cc, rr)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) blk[3 * i + j] = seg[9 * p + 3 * j + i];
#pragma unroll
        for (int q = 0; q < 9; ++q) seg[9 * s2 + q] = blk[q];
      } else {
#pragma unroll
        for (int q = 0; q < 9; ++q) blk[q] = seg[9 * s2 + q];
      }
#pragma unroll
      for (int q = 0; q < 9; ++q) sum[q] += blk[q];
      if constexpr (WITH_R) {
        const double du0 = A.u[3 * (size_t)cc] - ur[0], du1 = A.u[3 * (size_t)cc + 1] - ur[1],
                     du2 = A.u[3 * (size_t)cc + 2] - ur[2];
#pragma unroll
        for (int i = 0; i < 3; ++i) ru[i] += blk[3 * i] * du0 + blk[3 * i + 1] * du1 + blk[3 * i + 2] * du2;
      }
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) sum[q] = sum16(sum[q]);
    if constexpr (WITH_R) {
#pragma unroll
      for (int i = 0; i < 3; ++i) ru[i] = sum16(ru[i]);
    }
    dslot = max16(dslot);
    if (lane < 9 && dslot >= 0) {
      double v = sum[0];
#pragma unroll
      for (int q = 1; q < 9; ++q)
        if (lane == q) v = sum[q];
      seg[9 * dslot + lane] = -v;
    }
    if constexpr (WITH_R) {
      if (lane < 3) {
        const double v = lane == 0 ? ru[0] : (lane == 1 ? ru[1] : ru[2]);
        A.out_vec[3 * (size_t)rr + lane] = v - fl[3 * lr + lane];
      }
    }
  }
}

// Phase 4 (after the row barrier): stream the segment out.
template <int NT>
__device__ __forceinline__ void jac_stream(const ChunkArgs& A, int b0, int nb, const double* seg) {
  // Stream the segment out, 16 bytes per store where the destination allows, with
  // non-temporal stores: nothing in this kernel reads the values again, and bypassing
  // the caches measured 12 % faster on the 128x64x64 workload.
  typedef double v2d __attribute__((ext_vector_type(2)));
  double* dst = A.out_vals + (size_t)b0 * 9;
  const int n = 9 * nb;
  const int head = (int)(((uintptr_t)dst >> 3) & 1);  // one double before a 16-byte boundary
  if (head && threadIdx.x == 0 && n > 0) __builtin_nontemporal_store(seg[0], dst);
  const int n2 = (n - head) / 2;
  for (int k = threadIdx.x; k < n2; k += NT) {
    const int e = head + 2 * k;
    __builtin_nontemporal_store(v2d{seg[e], seg[e + 1]}, reinterpret_cast<v2d*>(dst + e));
  }
  if (((n - head) & 1) && threadIdx.x == 0) __builtin_nontemporal_store(seg[n - 1], dst + n - 1);
}

template <int BITS, bool WITH_R, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT >= 512 ? 4 : 1)))
void k_jac_chunk(ChunkArgs A) {
  extern __shared__ double lds[];
  const JacLds Lo = jac_lds_layout(A.max_rows, A.max_blocks, A.max_nodes);
  char* base = reinterpret_cast<char*>(lds);
  double* seg = reinterpret_cast<double*>(base + Lo.seg);
  double4* nodes = reinterpret_cast<double4*>(base + Lo.nodes);
  int* roff = reinterpret_cast<int*>(base + Lo.roff);
  int* col = reinterpret_cast<int*>(base + Lo.col);
  double* fl = reinterpret_cast<double*>(base + Lo.fl);
  const JacChunk J = jac_chunk_info(A, xcd_chunk(blockIdx.x, A.nchunks));
  jac_stage<NT, WITH_R>(A, J, seg, nodes, roff, col, fl);
  __syncthreads();
  jac_elements<BITS, NT, WITH_R>(A, J, seg, nodes, roff, fl);
  __syncthreads();
  jac_fixed_to_double<NT>(seg, 9 * J.nb, fixed_scale(A, J.c));
  __syncthreads();
  jac_rows<NT, WITH_R>(A, J, seg, roff, col, fl);
  __syncthreads();
  jac_stream<NT>(A, J.b0, J.nb, seg);
}

size_t chunk_lds_bytes(int mode, int max_rows, int max_blocks, int max_nodes) {
  if (mode == kModeJacobian || mode == kModeJacobianResidual) return jac_lds_layout(max_rows, max_blocks, max_nodes).total;
  const int blkw = mode == kModeGradZT ? 3 : (mode == kModeGradX ? 9 : 0);
  const int roww = (mode == kModeResidual || mode == kModePNormGrad || mode == kModeEnergyGradX || mode == kModeVolume ||
                    mode == kModePNormGradX)
                       ? 3
                       : ((mode == kModeEnergyGradZ || mode == kModeAdjZ) ? 1 : 0);
  const int row2 = mode == kModeResidual ? 3 : ((mode == kModePNormGrad || mode == kModeVolume) ? 1 : 0);
  const int nodew = mode == kModeAdjZ ? 10 : 7;
  size_t bytes = sizeof(double) * (nodew * (size_t)max_nodes + (size_t)max_blocks * blkw +
                                   (size_t)max_rows * (roww + row2));
  if (blkw > 0) bytes += sizeof(int) * (size_t)(max_rows + 1);
  return bytes;
}

// Work-group size per mode: the LDS-heavy Jacobian modes use 512 threads so that two
// work-groups (16 waves) share a CU's 160 KiB of LDS.
constexpr int kNtJac = 512;
constexpr int kNtVec = 256;

template <int BITS>
static hipError_t launch_chunk_b(int mode, const ChunkArgs& a, hipStream_t s) {
  const size_t lds = chunk_lds_bytes(mode, a.max_rows, a.max_blocks, a.max_nodes);
  dim3 grid(a.nchunks);
  switch (mode) {
    case kModeJacobian: k_jac_chunk<BITS, false, kNtJac><<<grid, kNtJac, lds, s>>>(a); break;
    case kModeJacobianResidual: k_jac_chunk<BITS, true, kNtJac><<<grid, kNtJac, lds, s>>>(a); break;
    case kModeResidual: k_chunk<BITS, kModeResidual, kNtVec><<<grid, kNtVec, lds, s>>>(a); break;
    case kModeGradZT: k_chunk<BITS, kModeGradZT, kNtVec><<<grid, kNtVec, lds, s>>>(a); break;
    case kModeEnergy: k_chunk<BITS, kModeEnergy, kNtVec><<<grid, kNtVec, lds, s>>>(a); break;
    case kModeEnergyGradZ: k_chunk<BITS, kModeEnergyGradZ, kNtVec><<<grid, kNtVec, lds, s>>>(a); break;
    case kModePNorm: k_chunk<BITS, kModePNorm, kNtVec><<<grid, kNtVec, lds, s>>>(a); break;
    case kModePNormGrad: k_chunk<BITS, kModePNormGrad, kNtVec><<<grid, kNtVec, lds, s>>>(a); break;
    case kModeAdjZ: k_chunk<BITS, kModeAdjZ, kNtVec><<<grid, kNtVec, lds, s>>>(a); break;
    case kModeGradX: k_chunk<BITS, kModeGradX, kNtVec><<<grid, kNtVec, lds, s>>>(a); break;
    case kModeEnergyGradX: k_chunk<BITS, kModeEnergyGradX, kNtVec><<<grid, kNtVec, lds, s>>>(a); break;
    case kModeVolume: k_chunk<BITS, kModeVolume, kNtVec><<<grid, kNtVec, lds, s>>>(a); break;
    case kModePNormGradX: k_chunk<BITS, kModePNormGradX, kNtVec><<<grid, kNtVec, lds, s>>>(a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename K>
static void allow_lds(K kernel) {
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                            160 * 1024);
}

hipError_t launch_chunk(int bits, int mode, const ChunkArgs& a, hipStream_t s) {
  static bool once = false;
  if (!once) {
    once = true;
    allow_lds(&k_jac_chunk<4, false, kNtJac>);
    allow_lds(&k_jac_chunk<4, true, kNtJac>);
    allow_lds(&k_jac_chunk<8, false, kNtJac>);
    allow_lds(&k_jac_chunk<8, true, kNtJac>);
    allow_lds(&k_chunk<4, kModeGradZT, kNtVec>);
    allow_lds(&k_chunk<8, kModeGradZT, kNtVec>);
    allow_lds(&k_chunk<4, kModeGradX, kNtVec>);
    allow_lds(&k_chunk<8, kModeGradX, kNtVec>);
  }
  if (a.nchunks == 0) return hipSuccess;
  return bits == 4 ? launch_chunk_b<4>(mode, a, s) : launch_chunk_b<8>(mode, a, s);
}

}  // namespace pa
