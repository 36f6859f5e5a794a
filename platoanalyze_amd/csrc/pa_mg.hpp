// Geometric multigrid on box node grids: transfer operators, density homogenisation,
// Chebyshev block-Jacobi smoothing, the coarsest dense solve and the PCG vector steps
// of the multigrid-preconditioned solve.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pa_kernels.hpp"

namespace pa {

struct BoxDims {
  int nx, ny, nz;  // node counts per axis
};

// rc[c] = sum_{fine i within one fine spacing of coarse node c} w_i rf[i] with the
// full-weighting weights (1, 1/2 per off-axis step), i.e. R = P^T of trilinear P;
// constrained coarse dofs are set to 0.  3 components per node.
// When cx is given, the coarse level's pre-smoothing starts in the same kernel:
// cx = cd = D_c^-1 rc * inv_theta.
hipError_t launch_mg_restrict(const double* rf, double* rc, const int32_t* cgrid, const int32_t* fnode_of_grid,
                              BoxDims fd, const uint8_t* cbcmask, int ncn, hipStream_t s,
                              const double* cdinv = nullptr, double inv_theta = 0.0, double* cx = nullptr,
                              double* cd = nullptr);
// xf[i] += sum_parents w ec[parent] (trilinear), skipping constrained fine dofs.
hipError_t launch_mg_prolong_add(const double* ec, double* xf, const int32_t* fgrid, const int32_t* cnode_of_grid,
                                 BoxDims cd, const uint8_t* fbcmask, int nfn, hipStream_t s);
// Normalised full weighting of a scalar nodal field (density homogenisation through the
// SIMP stiffness weight): wc[c] = sum w_i wf[i] / sum w_i.
hipError_t launch_mg_wrestrict(const double* wf, double* wc, const int32_t* cgrid, const int32_t* fnode_of_grid,
                               BoxDims fd, int ncn, hipStream_t s);
// nodal stiffness weight w = w(z) of the material's penalty model (coarse levels use w
// as their density under the identity penalty)
hipError_t launch_simp_w(const double* z, double* w, int n, const Material& m, hipStream_t s);

// Chebyshev smoothing of K x = r with block-Jacobi D^-1, from x = 0:
//   first: d = D^-1 r / theta; x = d  (r - t in place of r when t is given; d not
//          written when null)
//   next:  d = c1 d + c2 D^-1 (r - Kx); x += d
hipError_t launch_cheb_first(double* x, double* d, const double* r, const double* dinv, double inv_theta, int nn,
                             hipStream_t s, const double* t = nullptr);
hipError_t launch_cheb_next(double* x, double* d, const double* r, const double* Kx, const double* dinv, double c1,
                            double c2, int nn, hipStream_t s);
// out = a - b ; y += x ; y = D^-1 x
hipError_t launch_sub(double* out, const double* a, const double* b, size_t n, hipStream_t s);
hipError_t launch_add(double* y, const double* x, size_t n, hipStream_t s);
hipError_t launch_add2(double* out, const double* a, const double* b, size_t n, hipStream_t s);  // out = a + b
hipError_t launch_dinv_apply(double* y, const double* dinv, const double* x, int nn, hipStream_t s);
// y = A x with A dense n x n (row major)
hipError_t launch_dense_gemv(const double* A, const double* x, double* y, int n, hipStream_t s);

// Smoothing steps fused with the block-CRS product K x (16 lanes per block row):
//   kMgRes:      out = r - K x
//   kMgResCheb0: out = r - K x;  y = d = D^-1 out * inv_theta   (post-smoothing start)
//   kMgCheb:     d = c1 d + c2 D^-1 (r - K x);  out = x + d     (out must not alias x)
//   kMgChebAdd:  d = c1 d + c2 D^-1 (r - K x);  out = base + x + d
enum MgRowMode { kMgRes = 0, kMgResCheb0 = 1, kMgCheb = 2, kMgChebAdd = 3 };
struct MgRowArgs {
  const double* vals;
  const float* vals32;  // when set: an FP32 copy of the level's matrix, read in place of vals
  const int32_t* rowmap;
  const int32_t* col;
  const double* x;
  const double* r;
  const double* dinv;
  double* d;
  double* out;
  double* y;
  const double* base;
  double c1, c2, inv_theta;
  int nrows;
};
hipError_t launch_mg_rows(int mode, const MgRowArgs& a, hipStream_t s);
// b = (float) a, n values
hipError_t launch_to_f32(const double* a, float* b, size_t n, hipStream_t s);
// y *= 1 / sqrt(scal[slot])  (power-iteration normalisation without a host round trip)
hipError_t launch_scale_rsqrt(double* y, size_t n, const double* scal, int slot, hipStream_t s);

// PCG steps of the multigrid-preconditioned solve (scalars in scal, see ScalarSlot):
//   update: x += alpha p; r -= alpha q (alpha = scal[kAlpha]); partials of r.r
//   pupdate: p = z + beta p (beta = scal[kBeta])
hipError_t launch_mg_update(double* x, double* r, const double* p, const double* q, const double* scal, int nn,
                            double* P_rr, int grid, hipStream_t s);
hipError_t launch_mg_pupdate(double* p, const double* z, const double* scal, size_t n, hipStream_t s);

}  // namespace pa
