// Mixed-precision operators of the Tet4 elastostatics path: an FP32 matrix-free K p
// inside an FP64 Jacobi-PCG (FP64 vectors, reductions and block-diagonal
// preconditioner; residual replaced by the true FP64 residual f - K u).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pa_kernels.hpp"
#include "pa_reduce.hpp"

namespace pa {

struct MfArgs {
  // chunk structure (shared with the assembly kernels)
  const int32_t* chunk_rows;
  const int32_t* node_offs;
  const int32_t* nodes;
  const uint16_t* ent_ln;
  const int32_t* list_offs;
  int nchunks, max_rows, max_nodes;
  const float4* Xz;       // FP32 x, y, z, density per node
  const uint8_t* bcmask;  // bit i set: dof 3n+i constrained
  // direction update fused into the apply: p_cur = z + beta p_prev with beta =
  // scal[kBeta] (p_cur = z when first); FP64 vectors
  const double* z;
  const double* p_prev;
  double* p_cur;
  int first;
  const double* scal;
  double* y;         // y = K p_cur through the FP32 operator (constrained rows: y = p_cur)
  double* partials;  // per-chunk partial of p_cur . y
  Material m;
  const float* gbound;  // per chunk: bound of vol |grad N|^2 (scale of the fixed-point sums)
};
hipError_t launch_mf_apply(const MfArgs& a, hipStream_t s);

// FP64 3x3 diagonal blocks of K(z) per node, inverted (constrained dofs -> identity).
struct DiagArgs {
  const int32_t* chunk_rows;
  const int32_t* node_offs;
  const int32_t* nodes;
  const uint16_t* ent_ln;
  const int32_t* list_offs;
  int nchunks, max_rows, max_nodes;
  const double* X;
  const double* z;
  const uint8_t* bcmask;
  double* dinv;
  Material m;
  const float* gbound;  // per chunk: bound of vol |grad N|^2 (scale of the fixed-point sums)
};
hipError_t launch_mf_diag_inv(const DiagArgs& a, hipStream_t s);

// PCG steps.  Vector kernels write per-work-group partials; one-work-group finishers
// turn them into the scalars of the next step.
// z = Dinv r, partials of r.z into P_rz.
hipError_t launch_pcgm_init(const double* r, double* z, const double* dinv, int nnodes, double* P_rz, int grid,
                            hipStream_t s);
// pq = sum of the n apply partials; scal[kAlpha] = scal[kRZ] / pq
hipError_t launch_pcgm_alpha(const double* partials, int n, double* scal, hipStream_t s);
// x += alpha p; r -= alpha q; z = Dinv r (alpha = scal[kAlpha]); partials of r.z, r.r
hipError_t launch_pcgm_update(double* x, double* r, double* z, const double* p, const double* q, const double* dinv,
                              int nnodes, const double* scal, double* P_rz, double* P_rr, int grid, hipStream_t s);
// rz_new = sum P_rz; scal[kRR] = sum P_rr (when P_rr is given), then by mode:
//   kBetaInit:    scal[kRZ] = rz_new
//   kBetaStep:    scal[kBeta] = rz_new / scal[kRZ]; scal[kRZPrev] = scal[kRZ]; scal[kRZ] = rz_new
//   kBetaReplace: scal[kBeta] = rz_new / scal[kRZPrev]; scal[kRZ] = rz_new  (rz_new of a
//                 replaced residual)
enum BetaMode { kBetaStep = 0, kBetaInit = 1, kBetaReplace = 2 };
hipError_t launch_pcgm_beta(const double* P_rz, const double* P_rr, int n, double* scal, int mode, hipStream_t s);
// r = mask(b - R) with b optional (R = K u - f from the residual kernel, or K u);
// partials of ||r||^2 into P, one per work-group of `grid`
hipError_t launch_neg_mask(const double* R, const double* b, const uint8_t* bcmask, double* r, int nnodes, double* P,
                           int grid, hipStream_t s);
// u -= omega Dinv (K u) at the free dofs (Ku from the residual kernel without load)
hipError_t launch_jacobi_step(double* u, const double* Ku, const double* dinv, const uint8_t* bcmask, int nnodes,
                              double omega, hipStream_t s);
// Xz[n] = (float) (X[3n], X[3n+1], X[3n+2], z[n])
hipError_t launch_make_xz(const double* X, const double* z, float4* Xz, int nnodes, hipStream_t s);

}  // namespace pa
