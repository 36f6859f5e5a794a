// Device kernels of the Tet4 SIMP elastostatics path (gfx950).
//
// Element math follows the reference functors:
//   gradients/volume   ComputeGradientWorkset<3>  src/ImplicitFunctors.hpp:165-217
//   strain (Voigt)     Strain<3>                  src/Strain.hpp:55-95
//   stress             LinearStress<3>            src/LinearStress.hpp:60-77
//   MSIMP weighting    ApplyWeighting + MSIMP     src/ApplyWeighting.hpp:49-65, src/Simp.hpp:130-134
//   divergence         StressDivergence<3>        src/StressDivergence.hpp:40-89
//   energy             ScalarProduct<6>           src/ScalarProduct.hpp:28-42
// The Jacobian block of an isotropic element is written in closed form,
//   K_ab = vol*w*(lam g_a g_b^T + mu g_b g_a^T + mu (g_a.g_b) I),
// which is what the reference's SFad<double,12> evaluation of the same residual yields.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pa {

constexpr int kThreads = 256;
constexpr int kNumXcd = 8;

// Penalty models of the stiffness weight w(rho) (ApplyWeighting, src/ApplyWeighting.hpp):
// MSIMP (src/Simp.hpp:130-134), RAMP (src/Ramp.hpp), Heaviside (src/Heaviside.hpp),
// NoPenalty (src/NoPenalty.hpp).
enum PenaltyModel { kPenaltySimp = 0, kPenaltyRamp = 1, kPenaltyHeaviside = 2, kPenaltyNone = 3 };

struct Material {
  double c00, c01, c33;  // Voigt stiffness: c(1-nu), c*nu, c(1-2nu)/2
  double p, rmin;        // penalty exponent and minimum value
  int pint;              // p as a small positive integer (1..8), else 0
  int model;             // PenaltyModel
  double len;            // Heaviside regularization length
};

// integer exponents by multiplication (uniform branch on a kernel argument), else pow()
__device__ __forceinline__ double simp_pow(double rho, double p, int pint) {
  if (pint > 0) {
    double r = rho;
    for (int k = 1; k < pint; ++k) r *= rho;
    return r;
  }
  return pow(rho, p);
}

// w(rho) and dw/drho of the material's penalty model
__device__ __forceinline__ double penalty_w(const Material& m, double rho) {
  switch (m.model) {
    case kPenaltyRamp: return m.rmin + (1.0 - m.rmin) * rho / (1.0 + m.p * (1.0 - rho));
    case kPenaltyHeaviside: {
      if (rho <= -m.len) return m.rmin;
      if (rho >= m.len) return 1.0;
      const double h = 0.5 * (1.0 + sin(M_PI * rho / (2.0 * m.len)));
      return m.rmin + (1.0 - m.rmin) * pow(h, m.p);
    }
    case kPenaltyNone: return 1.0;
    default: return m.rmin + (1.0 - m.rmin) * simp_pow(rho, m.p, m.pint);
  }
}
// Inverse of a 3x3 block (row-major) by cofactors, with the constrained dofs (bits of bm)
// replaced by identity rows and columns first: the block-Jacobi preconditioner of the
// constrained system.
__device__ __forceinline__ void block_inv3_masked(double a[9], uint32_t bm, double* __restrict__ o) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      if (((bm >> i) | (bm >> j)) & 1u) a[3 * i + j] = (i == j) ? 1.0 : 0.0;
  const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
  const double id = 1.0 / (a[0] * c00 + a[1] * c01 + a[2] * c02);
  o[0] = c00 * id;
  o[1] = (a[2] * a[7] - a[1] * a[8]) * id;
  o[2] = (a[1] * a[5] - a[2] * a[4]) * id;
  o[3] = c01 * id;
  o[4] = (a[0] * a[8] - a[2] * a[6]) * id;
  o[5] = (a[2] * a[3] - a[0] * a[5]) * id;
  o[6] = c02 * id;
  o[7] = (a[1] * a[6] - a[0] * a[7]) * id;
  o[8] = (a[0] * a[4] - a[1] * a[3]) * id;
}

__device__ __forceinline__ double penalty_dw(const Material& m, double rho) {
  switch (m.model) {
    case kPenaltyRamp: {
      const double d = 1.0 + m.p * (1.0 - rho);
      return (1.0 - m.rmin) * (1.0 + m.p) / (d * d);
    }
    case kPenaltyHeaviside: {
      if (rho <= -m.len || rho >= m.len) return 0.0;
      const double a = M_PI / (2.0 * m.len);
      const double h = 0.5 * (1.0 + sin(a * rho));
      return (1.0 - m.rmin) * m.p * pow(h, m.p - 1.0) * 0.5 * a * cos(a * rho);
    }
    case kPenaltyNone: return 0.0;
    default: return (1.0 - m.rmin) * m.p * simp_pow(rho, m.p - 1.0, m.pint - 1);
  }
}

// Arguments of the chunked element kernels (one work-group per row chunk).
struct ChunkArgs {
  const int4* conn;
  const uint32_t* codes;
  const double* X;
  const double* z;
  const double* u;
  const double* u2;   // second state (adjoint lambda) for kModeAdjZ
  const int32_t* rowmap;
  const int32_t* colidx;
  const int32_t* chunk_rows;
  const int32_t* list_offs;
  const uint32_t* list;
  const int32_t* node_offs;   // [nchunks+1] node-table offsets
  const int32_t* nodes;       // node tables (chunk rows first)
  const uint16_t* ent_ln;     // 4 local node indices per entry
  const uint32_t* ent_code;   // slot codes per entry
  const uint16_t* tslot;      // per graph block: chunk-local slot of the transposed block
  double* out_vals;   // Jacobian values (9/block) or dR/dz (3/block)
  double* out_vec;    // residual (3/row) or nodal gradient (1/row)
  double* out_vec2;   // p-norm dF/dz partial (1/row)
  double* partials;   // per-chunk partial sums (energy)
  // traction load of the chunk's rows, subtracted from the residual (R = K u - f)
  const int32_t* load_offs;  // [nchunks+1]
  const int32_t* load_row;   // row index relative to the chunk's first row
  const double* load_f;      // 3 per entry
  int sub_load;
  // per chunk: max over its elements of vol |grad N|_max^2 (rounded up), which bounds the
  // magnitude of every element block entry by gbound (|lambda| + 2 mu) for weights <= 1
  const float* gbound;
  int nchunks;
  int max_rows;
  int max_blocks;
  int max_nodes;
  // stress p-norm F = (sum vol m^p)^(1/p): measure 0 = Voigt norm of sigma_w (the
  // reference StressPNorm), 1 = relaxed von Mises rho^q vm(sigma0); pscale multiplies
  // the gradients (pow(S,(1-p)/p)/p, known after the value pass)
  int measure;
  double pn, q, pscale;
  Material m;
};

enum ChunkMode { kModeJacobian = 0, kModeJacobianResidual = 1, kModeResidual = 2, kModeGradZT = 3,
                 kModeEnergy = 4, kModeEnergyGradZ = 5, kModePNorm = 6, kModePNormGrad = 7, kModeAdjZ = 8,
                 kModeGradX = 9, kModeEnergyGradX = 10, kModeVolume = 11, kModePNormGradX = 12,
                 kNumChunkModes = 13 };


// launch helpers (defined in pa_kernels.hip)
size_t chunk_lds_bytes(int mode, int max_rows, int max_blocks, int max_nodes);
hipError_t launch_chunk(int bits, int mode, const ChunkArgs& a, hipStream_t s);

}  // namespace pa
