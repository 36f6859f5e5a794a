// Deflated PCG (Saad, Yeung, Erhel and Guyomarc'h, SIAM J. Sci. Comput. 21(5), 2000):
// with k vectors W and E = W^T A W, the start is corrected by W E^-1 W^T r0 and every
// search direction is projected A-orthogonally to span(W): p -= W E^-1 (A W)^T z.  The
// kernels below evaluate the k dot products (fixed-order partials), the small solve and
// the k-term update on the device, so an iteration needs no host round trip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace pa {

constexpr int kMaxDeflation = 16;

// P[j * gridDim.x + block] = partial of V_j . x over the block's nodes (3 dofs each); the
// columns of V are ld apart.  Persistent XCD-contiguous grid (red_grid(nnodes, 256)).
__global__ void k_multi_dot(const double* V, int k, size_t ld, const double* x, int nnodes, double* P);

// mu = Einv (sum of the np partials of each column); Einv k x k row-major.  One work-group.
__global__ void k_defl_mu(const double* P, int np, int k, const double* Einv, double* mu);

// y += sign * sum_j mu_j V_j over n entries.
__global__ void k_multi_axpy(double* y, const double* V, int k, size_t ld, const double* mu, double sign, size_t n);

}  // namespace pa
