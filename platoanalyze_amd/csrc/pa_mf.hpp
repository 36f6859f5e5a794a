// Matrix-free operators of the Tet4 elastostatics path (FP32 apply for the inner PCG
// of the mixed-precision solve, FP64 block diagonal for its Jacobi preconditioner).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pa_kernels.hpp"

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
  const float* x;         // input vector (3 per node)
  float* y;               // output y = K x (constrained rows: y = x)
  double* partials;       // per-chunk partial of x.y (FP64)
  unsigned* counter;      // last-block counter
  double* scal;           // scalar slots (see ScalarSlot)
  int scal_out;           // slot receiving x.y (or -1)
  int scal_num;           // alpha = scal[scal_num] / x.y into kAlpha when >= 0
  Material m;
};

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

// FP32 PCG vector steps with FP64 reductions (mirror of the FP64 solver kernels).
hipError_t launch_pcg32_init(const float* r, float* z, float* p, const float* dinv, int nnodes, double* partials,
                             unsigned* counter, double* scal, hipStream_t s);
hipError_t launch_pcg32_update(float* x, float* r, float* z, const float* p, const float* q, const float* dinv,
                               int nnodes, double* partials, unsigned* counter, double* scal, hipStream_t s);
hipError_t launch_pcg32_pupdate(float* p, const float* z, int ndof, const double* scal, hipStream_t s);

// r32 = mask(b - R) with b optional (R = K u - f from the residual kernel, or K x),
// ||r||^2 into scal[slot]
hipError_t launch_neg_mask_to_f32(const double* R, const double* b, const uint8_t* bcmask, float* r32, int nnodes,
                                  double* partials, unsigned* counter, double* scal, int slot, hipStream_t s);
// Xz[n] = (float) (X[3n], X[3n+1], X[3n+2], z[n])
hipError_t launch_make_xz(const double* X, const double* z, float4* Xz, int nnodes, hipStream_t s);
// u += d (FP64 += FP32)
hipError_t launch_axpy_f32(double* u, const float* d, size_t n, hipStream_t s);

}  // namespace pa
