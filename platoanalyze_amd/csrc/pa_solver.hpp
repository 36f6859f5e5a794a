// Block-CRS solver kernels (declarations); see pa_solver.hip.
#pragma once
#include <hip/hip_runtime.h>

namespace pa {

// slots of the device scalar array used by the PCG kernels
enum ScalarSlot { kRZ = 0, kPQ = 1, kAlpha = 2, kBeta = 3, kRR = 4, kBB = 5, kDot0 = 6, kDot1 = 7,
                  kEnergy = 8, kRZPrev = 9, kNumScal = 16 };

// partial arrays of the solver (kRedBlocks doubles each) in one device buffer
enum PartialSlot { kPPQ = 0, kPRZ0 = 1, kPRZ1 = 2, kPRR = 3, kPBB = 4, kPDot = 5, kNumPartialSlots = 6 };

__global__ void k_spmv(const double* vals, const int* rowmap, const int* col, const double* x,
                       const double* b, double* y, int nrows, int mode);
template <int BR>
__global__ void k_matvec_add(const double* vals, const int* rowmap, const int* col, const double* x, double* y,
                             int nrows, double alpha);
__global__ void k_pcg_spmv(const double* vals, const int* rowmap, const int* col, const double* p, double* q,
                           int nrows, double* P_pq);
__global__ void k_pcg_pupdate(double* p, const double* z, int nnodes, const double* P_rz_new,
                              const double* P_rz_old, int np);
__global__ void k_pcg_update(double* x, double* r, double* z, const double* p, const double* q,
                             const double* dinv, int nnodes, const double* P_pq, int np_pq, const double* P_rz,
                             int np, double* P_rz_next, double* P_rr, double* scal);
__global__ void k_pcg_init(const double* r, double* z, double* p, const double* b, const double* dinv,
                           int nnodes, double* P_rz, double* P_rr, double* P_bb);
__global__ void k_dot(const double* a, const double* b, size_t n, double* P);
__global__ void k_sum(const double* partials, int n, double* scal, int slot);
__global__ void k_diag_inv(const double* vals, const int* diag_slot, double* dinv, int nnodes);
__global__ void k_add_load(double* y, const int* nodes, const double* f, int n, double s);
__global__ void k_fill(double* y, size_t n, double v);
// y[idx[k]] += v[k] (distinct positions)
__global__ void k_scatter_add(double* y, const int64_t* idx, const double* v, size_t n);
__global__ void k_scale(double* y, size_t n, double s);
__global__ void k_axpy(double* y, const double* x, size_t n, double a);
__global__ void k_bc_lift(const double* vals, const int* rowmap, const int* col, const int* dofs,
                          const double* bcv, int nbc, double scale, double* rhs);
__global__ void k_bc_zero(double* vals, const int* rowmap, const int* col, const int* dofs, int nbc);
__global__ void k_bc_set(double* rhs, const int* dofs, const double* bcv, int nbc, double scale);
__global__ void k_bc_zero_vec(double* y, const int* dofs, int nbc);

}  // namespace pa
