// Matrix-free operators of the Tet4 elastostatics path (FP32 apply for the inner PCG
// of the mixed-precision solve, FP64 block diagonal for its Jacobi preconditioner).
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
  // data
  const float4* Xz;       // FP32 x, y, z, density per node
  const uint8_t* bcmask;  // bit i set: dof 3n+i constrained
  // direction update fused into the apply: p_cur = z + beta p_prev (p_cur = z when
  // first), beta = sum(P_rz_cur) / sum(P_rz_prev) over np partials
  const float* z;
  const float* p_prev;
  float* p_cur;
  const double* P_rz_cur;
  const double* P_rz_prev;
  int np, first;
  float* y;     // y = K p_cur (constrained rows: y = p_cur)
  double* P_pq; // per-work-group partial of p_cur . y (FP64), gridDim.x of them
  Material m;
};

// grid of the persistent matrix-free apply (multiple of 8, <= kRedBlocks)
int mf_apply_grid(int nchunks);
hipError_t launch_mf_apply(const MfArgs& a, hipStream_t s);

// FP64 3x3 diagonal blocks of K(z) per node, inverted (constrained dofs -> identity),
// written as FP32 (9 per node).
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
  float* dinv;
  Material m;
};
hipError_t launch_mf_diag_inv(const DiagArgs& a, hipStream_t s);

// FP32 PCG vector steps with FP64 partials (the FP64 solver's protocol: each kernel
// writes per-work-group partials that the next kernel sums in a fixed order).
// z = Dinv r, partials of r.z into P_rz.
hipError_t launch_pcg32_init(const float* r, float* z, const float* dinv, int nnodes, double* P_rz, int grid,
                             hipStream_t s);
// alpha = sum(P_pq)/sum(P_rz); x += alpha p; r -= alpha q; z = Dinv r; partials of
// r.z and r.r.
hipError_t launch_pcg32_update(float* x, float* r, float* z, const float* p, const float* q, const float* dinv,
                               int nnodes, const double* P_pq, int np_pq, const double* P_rz, int np,
                               double* P_rz_next, double* P_rr, double* scal, int grid, hipStream_t s);

// r32 = mask(b - R) with b optional (R = K u - f from the residual kernel, or K x);
// partials of ||r||^2 (FP64) into P, one per work-group of `grid`
hipError_t launch_neg_mask_to_f32(const double* R, const double* b, const uint8_t* bcmask, float* r32, int nnodes,
                                  double* P, int grid, hipStream_t s);
// Xz[n] = (float) (X[3n], X[3n+1], X[3n+2], z[n])
hipError_t launch_make_xz(const double* X, const double* z, float4* Xz, int nnodes, hipStream_t s);
// y = 0 (FP32)
hipError_t launch_zero_f32(float* y, size_t n, hipStream_t s);
// u += d (FP64 += FP32)
hipError_t launch_axpy_f32(double* u, const float* d, size_t n, hipStream_t s);

}  // namespace pa
