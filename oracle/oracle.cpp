// ORACLE -- test infrastructure only.  A CPU restatement of the reference's Tet4
// elastostatics path used as the checker by tests/, __graft_entry__.smoke() and the
// CPU-baseline leg of bench.py.  Nothing in the product library links or calls it.
//
// It follows the reference algorithm step by step, including its cost structure:
// per-cell worksets, forward-mode dual numbers with 12 (state) and 4 (control)
// derivative slots (Sacado SFad, src/SimplexFadTypes.hpp:13-24), residual and
// Jacobian through the same element functors, atomic scatter after a linear search
// of the CRS row, a second all-cells "boundary" pass, Dirichlet elimination, and a
// Jacobi-preconditioned CG.
//
// Pinning (what checks each function; details in tests/test_oracle_golden.py):
//   orc_box / node numbering ......... DerivativeTests.cpp residual (81), grad_z (27),
//                                      value 46.125; OmegaHUtilitiesTests LocalElementCoords_3D
//   element functors ................. DerivativeTests.cpp:159-253 (gradients, strain,
//                                      stress, element residual of six cells)
//   orc_residual ..................... DerivativeTests.cpp:361-398 (81 values)
//   orc_jacobian ..................... DerivativeTests.cpp:408-428 (45 entries), symmetry,
//                                      rigid-body null space, finite differences of orc_residual
//   orc_gradient_z ................... DerivativeTests.cpp:444-453 (24 entries)
//   orc_energy ....................... DerivativeTests.cpp:585-652 (value, grad_u, grad_z)
//   orc_graph ........................ CreateBlockMatrix rule; row lengths = vertex degree + 1
//   orc_apply_constraints / orc_pcg .. dense LAPACK solve on small meshes (numpy)
//   orc_traction ..................... total force = area * traction (closed form)
//   orc_stress_pnorm / orc_pnorm ..... DerivativeTests.cpp:1058 (14525.25169157 at p=12),
//                                      :1069-1108 (grad_u 27, grad_z 27)
//   orc_pnorm_gradient_x ............. DerivativeTests.cpp:1113-1137 (27 values at p=12)
//   vm_of / orc_von_mises ............ PlatoAugLagStressTest.cpp:1063-1101 (two cells)
//   orc_gradient_x ................... DerivativeTests.cpp:469-483 (36 entries); central
//                                      differences of orc_residual in the coordinates
//   orc_energy_gradient_x ............ DerivativeTests.cpp:665-693 (81 values); central
//                                      differences of orc_energy in the coordinates
//   orc_volume ....................... DerivativeTests.cpp:2151-2302 (value 1, dV/dz 27,
//                                      dV/dX 81) with linear SIMP; closed forms per model
//   penalty weights at interior densities (SIMP p=3 rmin=1e-3, RAMP, Heaviside, none):
//                                      hand-computed w(c) of src/Simp.hpp:130-134,
//                                      src/Ramp.hpp:94-100, src/Heaviside.hpp:110-125
//                                      against K(z=c) = w(c) K(1) on the whole mesh and
//                                      K(z) = w(mean z) K(1) on one element
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

namespace {

// ----------------------------------------------------------------- dual numbers
template <int N>
struct Fad {
  double v;
  double dx[N];
  Fad() : v(0.0) {
    for (int i = 0; i < N; ++i) dx[i] = 0.0;
  }
  Fad(double x) : v(x) {
    for (int i = 0; i < N; ++i) dx[i] = 0.0;
  }
  Fad(double x, int k) : v(x) {  // seeded: d/dx_k = 1
    for (int i = 0; i < N; ++i) dx[i] = 0.0;
    dx[k] = 1.0;
  }
  Fad& operator+=(const Fad& o) {
    v += o.v;
    for (int i = 0; i < N; ++i) dx[i] += o.dx[i];
    return *this;
  }
  Fad& operator*=(double s) {
    v *= s;
    for (int i = 0; i < N; ++i) dx[i] *= s;
    return *this;
  }
};
template <int N>
inline Fad<N> operator+(const Fad<N>& a, const Fad<N>& b) {
  Fad<N> r(a.v + b.v);
  for (int i = 0; i < N; ++i) r.dx[i] = a.dx[i] + b.dx[i];
  return r;
}
template <int N>
inline Fad<N> operator-(const Fad<N>& a, const Fad<N>& b) {
  Fad<N> r(a.v - b.v);
  for (int i = 0; i < N; ++i) r.dx[i] = a.dx[i] - b.dx[i];
  return r;
}
template <int N>
inline Fad<N> operator*(const Fad<N>& a, const Fad<N>& b) {
  Fad<N> r(a.v * b.v);
  for (int i = 0; i < N; ++i) r.dx[i] = a.dx[i] * b.v + a.v * b.dx[i];
  return r;
}
template <int N>
inline Fad<N> operator*(const Fad<N>& a, double b) {
  Fad<N> r(a.v * b);
  for (int i = 0; i < N; ++i) r.dx[i] = a.dx[i] * b;
  return r;
}
template <int N>
inline Fad<N> operator*(double a, const Fad<N>& b) {
  return b * a;
}
template <int N>
inline Fad<N> operator/(const Fad<N>& a, double b) {
  Fad<N> r(a.v / b);
  for (int i = 0; i < N; ++i) r.dx[i] = a.dx[i] / b;
  return r;
}
template <int N>
inline Fad<N> operator+(double a, const Fad<N>& b) {
  Fad<N> r(b);
  r.v = a + b.v;
  return r;
}
template <int N>
inline Fad<N> fad_pow(const Fad<N>& a, double c) {
  Fad<N> r(std::pow(a.v, c));
  const double d = (c == 0.0) ? 0.0 : c * std::pow(a.v, c - 1.0);
  for (int i = 0; i < N; ++i) r.dx[i] = a.dx[i] * d;
  return r;
}
inline double fad_pow(double a, double c) { return std::pow(a, c); }
template <int N>
inline Fad<N> fad_sqrt(const Fad<N>& a) {
  Fad<N> r(std::sqrt(a.v));
  const double d = a.v > 0.0 ? 0.5 / r.v : 0.0;
  for (int i = 0; i < N; ++i) r.dx[i] = a.dx[i] * d;
  return r;
}
inline double fad_sqrt(double a) { return std::sqrt(a); }
template <int N>
inline Fad<N> operator*(const Fad<N>& a, int b) { return a * (double)b; }
template <int N>
inline Fad<N> operator/(const Fad<N>& a, const Fad<N>& b) {
  Fad<N> r(a.v / b.v);
  for (int i = 0; i < N; ++i) r.dx[i] = (a.dx[i] * b.v - a.v * b.dx[i]) / (b.v * b.v);
  return r;
}
template <int N>
inline Fad<N> fad_abs(const Fad<N>& a) { return a.v < 0.0 ? a * -1.0 : a; }
inline double fad_abs(double a) { return std::fabs(a); }
// product of a stress-typed and a weight-typed value into the result type
template <typename R, typename A, typename B>
inline R mix(const A& a, const B& b) { return R(a * b); }
template <>
inline Fad<12> mix<Fad<12>, Fad<12>, double>(const Fad<12>& a, const double& b) { return a * b; }
template <>
inline Fad<4> mix<Fad<4>, double, Fad<4>>(const double& a, const Fad<4>& b) { return b * a; }

template <typename T>
struct ValueOf {
  static double get(const T& x) { return x.v; }
};
template <>
struct ValueOf<double> {
  static double get(const double& x) { return x; }
};

struct Params {
  double c[6][6];
  double p, rmin;
  int model;   // 0 SIMP, 1 RAMP, 2 Heaviside, 3 none
  double len;  // Heaviside regularization length
};

// Penalty model for the following evaluations (set by orc_set_penalty_model)
int g_penalty_model = 0;
double g_penalty_len = 1.0;

void make_params(double E, double nu, double p, double rmin, Params& P) {
  // IsotropicLinearElasticMaterial<3> (src/IsotropicLinearElasticMaterial.cpp:87-127)
  std::memset(P.c, 0, sizeof(P.c));
  const double coef = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
  P.c[0][0] = coef * (1.0 - nu);
  P.c[0][1] = coef * nu;
  P.c[0][2] = coef * nu;
  P.c[1][0] = coef * nu;
  P.c[1][1] = coef * (1.0 - nu);
  P.c[1][2] = coef * nu;
  P.c[2][0] = coef * nu;
  P.c[2][1] = coef * nu;
  P.c[2][2] = coef * (1.0 - nu);
  P.c[3][3] = 1.0 / 2.0 * coef * (1.0 - 2.0 * nu);
  P.c[4][4] = 1.0 / 2.0 * coef * (1.0 - 2.0 * nu);
  P.c[5][5] = 1.0 / 2.0 * coef * (1.0 - 2.0 * nu);
  P.p = p;
  P.rmin = rmin;
  P.model = g_penalty_model;
  P.len = g_penalty_len;
}

// ComputeGradientWorkset<3> (src/ImplicitFunctors.hpp:165-217) with the cubature
// weight applied afterwards (src/elliptic/ElastostaticResidual.hpp:181-182).
void gradient_workset(const double* cfg /*4x3*/, double* grad /*4x3*/, double& vol) {
  auto C = [&](int n, int d) { return cfg[3 * n + d]; };
  double j11 = C(0, 0) - C(3, 0), j12 = C(1, 0) - C(3, 0), j13 = C(2, 0) - C(3, 0);
  double j21 = C(0, 1) - C(3, 1), j22 = C(1, 1) - C(3, 1), j23 = C(2, 1) - C(3, 1);
  double j31 = C(0, 2) - C(3, 2), j32 = C(1, 2) - C(3, 2), j33 = C(2, 2) - C(3, 2);
  double detj = j11 * j22 * j33 + j12 * j23 * j31 + j13 * j21 * j32 - j11 * j23 * j32 - j12 * j21 * j33 -
                j13 * j22 * j31;
  double i11 = (j22 * j33 - j23 * j32) / detj, i12 = (j13 * j32 - j12 * j33) / detj,
         i13 = (j12 * j23 - j13 * j22) / detj;
  double i21 = (j23 * j31 - j21 * j33) / detj, i22 = (j11 * j33 - j13 * j31) / detj,
         i23 = (j13 * j21 - j11 * j23) / detj;
  double i31 = (j21 * j32 - j22 * j31) / detj, i32 = (j12 * j31 - j11 * j32) / detj,
         i33 = (j11 * j22 - j12 * j21) / detj;
  vol = std::fabs(detj);
  grad[0] = i11; grad[1] = i12; grad[2] = i13;
  grad[3] = i21; grad[4] = i22; grad[5] = i23;
  grad[6] = i31; grad[7] = i32; grad[8] = i33;
  grad[9] = -(i11 + i21 + i31);
  grad[10] = -(i12 + i22 + i32);
  grad[11] = -(i13 + i23 + i33);
  const double cub = 1.0 / 2.0 / 3.0;  // LinearTetCubRuleDegreeOne<3>::initialize
  vol *= cub;
}

// Strain<3> (src/Strain.hpp:55-95)
template <typename S>
void strain(const S* u /*12*/, const double* g /*4x3*/, S* e /*6*/) {
  int t = 0;
  for (int d = 0; d < 3; ++d) {
    e[t] = S(0.0);
    for (int n = 0; n < 4; ++n) e[t] += u[3 * n + d] * g[3 * n + d];
    ++t;
  }
  for (int J = 2; J >= 1; --J)
    for (int I = J - 1; I >= 0; --I) {
      e[t] = S(0.0);
      for (int n = 0; n < 4; ++n) e[t] += (u[3 * n + J] * g[3 * n + I] + u[3 * n + I] * g[3 * n + J]);
      ++t;
    }
}

// LinearStress<3> (src/LinearStress.hpp:60-77), reference strain 0
template <typename S>
void stress(const Params& P, const S* e, S* s) {
  for (int i = 0; i < 6; ++i) {
    s[i] = S(0.0);
    for (int j = 0; j < 6; ++j) s[i] += e[j] * P.c[i][j];
  }
}

// Penalty functions w(rho) and dw/drho: RAMP (src/Ramp.hpp), Heaviside
// (src/Heaviside.hpp), NoPenalty (src/NoPenalty.hpp); SIMP below through fad_pow.
double penalty_value(const Params& P, double rho, double* dw) {
  switch (P.model) {
    case 1: {
      const double d = 1.0 + P.p * (1.0 - rho);
      *dw = (1.0 - P.rmin) * (1.0 + P.p) / (d * d);
      return P.rmin + (1.0 - P.rmin) * rho / d;
    }
    case 2: {
      if (rho <= -P.len) {
        *dw = 0.0;
        return P.rmin;
      }
      if (rho >= P.len) {
        *dw = 0.0;
        return 1.0;
      }
      const double a = M_PI / (2.0 * P.len), h = 0.5 * (1.0 + std::sin(a * rho));
      *dw = (1.0 - P.rmin) * P.p * std::pow(h, P.p - 1.0) * 0.5 * a * std::cos(a * rho);
      return P.rmin + (1.0 - P.rmin) * std::pow(h, P.p);
    }
    default:
      *dw = 0.0;
      return 1.0;
  }
}
inline double penalty_of(const Params& P, double rho) {
  double dw;
  return penalty_value(P, rho, &dw);
}
template <int N>
inline Fad<N> penalty_of(const Params& P, const Fad<N>& rho) {
  double dw;
  Fad<N> r(penalty_value(P, rho.v, &dw));
  for (int i = 0; i < N; ++i) r.dx[i] = dw * rho.dx[i];
  return r;
}

// ApplyWeighting<3,6,Penalty> (src/ApplyWeighting.hpp:49-65) with MSIMP
// (src/Simp.hpp:130-134) or the model chosen by orc_set_penalty_model
template <typename Z>
Z msimp_weight(const Params& P, const Z* z /*4*/) {
  Z rho(0.0);
  for (int n = 0; n < 4; ++n) rho += z[n];
  rho = rho / 4.0;
  if (P.model != 0) return penalty_of(P, rho);
  return P.rmin + (1.0 - P.rmin) * fad_pow(rho, P.p);
}

// StressDivergence<3> (src/StressDivergence.hpp:72-87)
static const int kVoigt[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};
template <typename R, typename S>
void divergence(const S* s, const double* g, double vol, R* r /*12, accumulated*/) {
  for (int i = 0; i < 3; ++i)
    for (int n = 0; n < 4; ++n)
      for (int j = 0; j < 3; ++j) r[3 * n + i] += vol * s[kVoigt[i][j]] * g[3 * n + j];
}

// BlockMatrixEntryOrdinal (src/ImplicitFunctors.hpp:645-681): linear row search.
inline int block_entry(const int* rowmap, const int* col, int row, int colnode) {
  for (int k = rowmap[row]; k < rowmap[row + 1]; ++k)
    if (col[k] == colnode) return k;
  return -1;
}

inline void atomic_add(double* p, double v) {
#pragma omp atomic
  *p += v;
}

// Configuration-typed element residual and energy for the shape derivatives
// (VectorFunction::gradient_x, src/elliptic/VectorFunction.hpp:317-408, and
// PhysicsScalarFunction::gradient_x, src/elliptic/PhysicsScalarFunction.hpp:293-355):
// the same gradient, strain, stress, weighting and divergence arithmetic with the
// cell coordinates as dual numbers (GradientX: ConfigFad = SFad<double,12>).
template <typename T>
void gradient_workset_t(const T* cfg /*4x3*/, T* grad /*4x3*/, T& vol) {
  auto C = [&](int n, int d) { return cfg[3 * n + d]; };
  T j11 = C(0, 0) - C(3, 0), j12 = C(1, 0) - C(3, 0), j13 = C(2, 0) - C(3, 0);
  T j21 = C(0, 1) - C(3, 1), j22 = C(1, 1) - C(3, 1), j23 = C(2, 1) - C(3, 1);
  T j31 = C(0, 2) - C(3, 2), j32 = C(1, 2) - C(3, 2), j33 = C(2, 2) - C(3, 2);
  T detj = j11 * j22 * j33 + j12 * j23 * j31 + j13 * j21 * j32 - j11 * j23 * j32 - j12 * j21 * j33 -
           j13 * j22 * j31;
  grad[0] = (j22 * j33 - j23 * j32) / detj;
  grad[1] = (j13 * j32 - j12 * j33) / detj;
  grad[2] = (j12 * j23 - j13 * j22) / detj;
  grad[3] = (j23 * j31 - j21 * j33) / detj;
  grad[4] = (j11 * j33 - j13 * j31) / detj;
  grad[5] = (j13 * j21 - j11 * j23) / detj;
  grad[6] = (j21 * j32 - j22 * j31) / detj;
  grad[7] = (j12 * j31 - j11 * j32) / detj;
  grad[8] = (j11 * j22 - j12 * j21) / detj;
  for (int d = 0; d < 3; ++d) grad[9 + d] = (grad[d] + grad[3 + d] + grad[6 + d]) * -1.0;
  vol = fad_abs(detj) * (1.0 / 2.0 / 3.0);
}

template <typename T>
void element_stress_t(const Params& P, const T* g, const double* u, double w, T* eps, T* sw) {
  int t = 0;
  for (int d = 0; d < 3; ++d) {
    eps[t] = T(0.0);
    for (int n = 0; n < 4; ++n) eps[t] += g[3 * n + d] * u[3 * n + d];
    ++t;
  }
  for (int J = 2; J >= 1; --J)
    for (int I = J - 1; I >= 0; --I) {
      eps[t] = T(0.0);
      for (int n = 0; n < 4; ++n) eps[t] += g[3 * n + I] * u[3 * n + J] + g[3 * n + J] * u[3 * n + I];
      ++t;
    }
  for (int i = 0; i < 6; ++i) {
    T si(0.0);
    for (int j = 0; j < 6; ++j) si += eps[j] * P.c[i][j];
    sw[i] = si * w;
  }
}

struct Mesh {
  int nn, ne;
  const double* X;
  const int* conn;
};

// worksetConfig (src/WorksetBase.hpp / src/Assembly.hpp:1035-1089)
std::vector<double> workset_config(const Mesh& m) {
  std::vector<double> ws(12 * (size_t)m.ne);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < m.ne; ++e)
    for (int n = 0; n < 4; ++n)
      for (int d = 0; d < 3; ++d) ws[12 * (size_t)e + 3 * n + d] = m.X[3 * (size_t)m.conn[4 * e + n] + d];
  return ws;
}

// SurfaceLoadIntegral (src/SurfaceLoadIntegral.hpp:100-157), scale -1: R -= area/3 t
void traction(const Mesh& m, int nfaces, const int* fv, const double* t, double* R) {
  const double cub = 1.0 / 2.0 / 3.0;
  for (int f = 0; f < nfaces; ++f) {
    const int* v = fv + 3 * f;
    double j0[3], j1[3];
    for (int d = 0; d < 3; ++d) {
      j0[d] = m.X[3 * (size_t)v[0] + d] - m.X[3 * (size_t)v[2] + d];
      j1[d] = m.X[3 * (size_t)v[1] + d] - m.X[3 * (size_t)v[2] + d];
    }
    double j23 = j0[1] * j1[2] - j0[2] * j1[1];
    double j31 = j0[2] * j1[0] - j0[0] * j1[2];
    double j12 = j0[0] * j1[1] - j0[1] * j1[0];
    double w = -1.0 * cub * std::sqrt(j23 * j23 + j31 * j31 + j12 * j12);
    for (int a = 0; a < 3; ++a)
      for (int d = 0; d < 3; ++d) R[3 * (size_t)v[a] + d] += w * t[d];
  }
}

}  // namespace

extern "C" {

void orc_set_penalty_model(int model, double len) {
  g_penalty_model = model;
  g_penalty_len = len;
}

int orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

// CreateBlockMatrix (src/ImplicitFunctors.hpp:735-808) over an ascending vertex star
// built from the element edges.  colidx==NULL: only rowmap and *nnz are produced.
int orc_graph(int nn, int ne, const int* conn, int* rowmap, int* colidx, int* nnz) {
  std::vector<std::vector<int>> star(nn);
  for (int e = 0; e < ne; ++e)
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b)
        if (a != b) star[conn[4 * e + a]].push_back(conn[4 * e + b]);
  std::vector<int> a2ab(nn + 1, 0);
  std::vector<int> ab2b;
  for (int r = 0; r < nn; ++r) {
    auto& s = star[r];
    std::sort(s.begin(), s.end());
    s.erase(std::unique(s.begin(), s.end()), s.end());
    a2ab[r + 1] = a2ab[r] + (int)s.size();
    ab2b.insert(ab2b.end(), s.begin(), s.end());
  }
  *nnz = (int)ab2b.size() + nn;
  for (int row = 0; row < nn; ++row) rowmap[row] = a2ab[row] + row;
  rowmap[nn] = *nnz;
  if (!colidx) return 0;
  for (int row = 0; row < nn; ++row) {
    int off = a2ab[row];
    int R0 = a2ab[row] + row, R1 = a2ab[row + 1] + row + 1;
    int n = R1 - R0;
    int i_oh = 0;
    for (int i = 0; i < n; i_oh++, i++) {
      bool ins = false;
      if (i_oh == i && i_oh + off >= a2ab[row + 1]) ins = true;
      else if (i_oh == i && ab2b[i_oh + off] > row) ins = true;
      if (ins) {
        colidx[R0 + i] = row;
        i_oh--;
      } else {
        colidx[R0 + i] = ab2b[i_oh + off];
      }
    }
  }
  return 0;
}

// Per-cell worksets of the first `ncells` cells (gradients 12, volume 1, strain 6,
// stress 6, element residual 12), as the reference's DerivativeTests "3D" test reads them.
void orc_cell_worksets(int nn, int ne, const double* X, const int* conn, const double* u, double E, double nu,
                       int ncells, double* grad, double* vol, double* eps, double* sig, double* res) {
  Mesh m{nn, ne, X, conn};
  Params P;
  make_params(E, nu, 1.0, 0.0, P);
  auto cfg = workset_config(m);
  for (int c = 0; c < ncells; ++c) {
    double ue[12];
    for (int n = 0; n < 4; ++n)
      for (int d = 0; d < 3; ++d) ue[3 * n + d] = u[3 * (size_t)conn[4 * c + n] + d];
    gradient_workset(&cfg[12 * (size_t)c], grad + 12 * c, vol[c]);
    strain(ue, grad + 12 * c, eps + 6 * c);
    stress(P, eps + 6 * c, sig + 6 * c);
    for (int k = 0; k < 12; ++k) res[12 * c + k] = 0.0;
    divergence(sig + 6 * c, grad + 12 * c, vol[c], res + 12 * c);
  }
}

// VectorFunction::value (src/elliptic/VectorFunction.hpp:231-313): worksets, the two
// ElastostaticResidual kernels, atomic assemble_residual, then the boundary pass.
void orc_residual(int nn, int ne, const double* X, const int* conn, const double* u, const double* z, double E,
                  double nu, double p, double rmin, int nfaces, const int* face_verts, const double* t,
                  double* R) {
  Mesh m{nn, ne, X, conn};
  Params P;
  make_params(E, nu, p, rmin, P);
  auto cfg = workset_config(m);
  std::vector<double> ustate(12 * (size_t)ne), zc(4 * (size_t)ne), grad(12 * (size_t)ne), vol(ne),
      sig(6 * (size_t)ne), res(12 * (size_t)ne, 0.0);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < ne; ++e)
    for (int n = 0; n < 4; ++n) {
      for (int d = 0; d < 3; ++d) ustate[12 * (size_t)e + 3 * n + d] = u[3 * (size_t)conn[4 * e + n] + d];
      zc[4 * (size_t)e + n] = z[conn[4 * e + n]];
    }
#pragma omp parallel for schedule(static)
  for (int e = 0; e < ne; ++e) {  // "Cauchy stress"
    gradient_workset(&cfg[12 * (size_t)e], &grad[12 * (size_t)e], vol[e]);
    double eps[6];
    strain(&ustate[12 * (size_t)e], &grad[12 * (size_t)e], eps);
    stress(P, eps, &sig[6 * (size_t)e]);
  }
#pragma omp parallel for schedule(static)
  for (int e = 0; e < ne; ++e) {  // "Apply weighting and compute divergence"
    double w = msimp_weight(P, &zc[4 * (size_t)e]);
    double sw[6];
    for (int k = 0; k < 6; ++k) sw[k] = sig[6 * (size_t)e + k] * w;
    divergence(sw, &grad[12 * (size_t)e], vol[e], &res[12 * (size_t)e]);
  }
  std::memset(R, 0, sizeof(double) * 3 * (size_t)nn);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < ne; ++e)  // assemble_residual (src/Assembly.hpp:1189-1246)
    for (int n = 0; n < 4; ++n)
      for (int d = 0; d < 3; ++d) atomic_add(&R[3 * (size_t)conn[4 * e + n] + d], res[12 * (size_t)e + 3 * n + d]);
  if (nfaces > 0) traction(m, nfaces, face_verts, t, R);
}

// VectorFunction::gradient_u (src/elliptic/VectorFunction.hpp:505-594) with
// JacobianTypes (StateFad = SFad<double,12>) and assemble_jacobian_fad
// (src/Assembly.hpp:1341-1402), including the all-cells boundary pass that adds the
// (zero) state derivative of the traction.  `values` is overwritten.
void orc_jacobian(int nn, int ne, const double* X, const int* conn, const double* u, const double* z, double E,
                  double nu, double p, double rmin, const int* rowmap, const int* col, double* values) {
  using F = Fad<12>;
  Mesh m{nn, ne, X, conn};
  Params P;
  make_params(E, nu, p, rmin, P);
  const size_t nnzv = 9 * (size_t)rowmap[nn];
  std::memset(values, 0, sizeof(double) * nnzv);
  auto cfg = workset_config(m);
  std::vector<F> ustate(12 * (size_t)ne), res(12 * (size_t)ne), sig(6 * (size_t)ne);
  std::vector<double> zc(4 * (size_t)ne), grad(12 * (size_t)ne), vol(ne);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < ne; ++e)  // workset_state_scalar_fad
    for (int n = 0; n < 4; ++n) {
      for (int d = 0; d < 3; ++d)
        ustate[12 * (size_t)e + 3 * n + d] = F(u ? u[3 * (size_t)conn[4 * e + n] + d] : 0.0, 3 * n + d);
      zc[4 * (size_t)e + n] = z[conn[4 * e + n]];
    }
#pragma omp parallel for schedule(static)
  for (int e = 0; e < ne; ++e) {
    gradient_workset(&cfg[12 * (size_t)e], &grad[12 * (size_t)e], vol[e]);
    F eps[6];
    strain(&ustate[12 * (size_t)e], &grad[12 * (size_t)e], eps);
    stress(P, eps, &sig[6 * (size_t)e]);
  }
#pragma omp parallel for schedule(static)
  for (int e = 0; e < ne; ++e) {
    double w = msimp_weight(P, &zc[4 * (size_t)e]);
    F sw[6];
    for (int k = 0; k < 6; ++k) sw[k] = sig[6 * (size_t)e + k] * w;
    divergence(sw, &grad[12 * (size_t)e], vol[e], &res[12 * (size_t)e]);
  }
  for (int pass = 0; pass < 2; ++pass) {
    // pass 1 = evaluate_boundary over all cells: traction has no state dependence,
    // its Fad result is zero and is scattered the same way.
#pragma omp parallel for schedule(static)
    for (int e = 0; e < ne; ++e) {
      for (int k = 0; k < 12; ++k) {
        const int rn = conn[4 * e + k / 3], id = k % 3;
        for (int l = 0; l < 12; ++l) {
          const int cn = conn[4 * e + l / 3], jd = l % 3;
          const int s = block_entry(rowmap, col, rn, cn);
          const double v = pass == 0 ? res[12 * (size_t)e + k].dx[l] : 0.0;
          atomic_add(&values[9 * (size_t)s + 3 * id + jd], v);
        }
      }
    }
  }
}

// VectorFunction::gradient_z (src/elliptic/VectorFunction.hpp:598-689): ControlFad =
// SFad<double,4>, transposed 1x3 block assembly (src/Assembly.hpp:1422-1485).
void orc_gradient_z(int nn, int ne, const double* X, const int* conn, const double* u, const double* z, double E,
                    double nu, double p, double rmin, const int* rowmap, const int* col, double* values_t) {
  using F = Fad<4>;
  Mesh m{nn, ne, X, conn};
  Params P;
  make_params(E, nu, p, rmin, P);
  std::memset(values_t, 0, sizeof(double) * 3 * (size_t)rowmap[nn]);
  auto cfg = workset_config(m);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < ne; ++e) {
    double g[12], v, ue[12];
    gradient_workset(&cfg[12 * (size_t)e], g, v);
    F zc[4];
    for (int n = 0; n < 4; ++n) {
      zc[n] = F(z[conn[4 * e + n]], n);
      for (int d = 0; d < 3; ++d) ue[3 * n + d] = u[3 * (size_t)conn[4 * e + n] + d];
    }
    double eps[6], s[6];
    strain(ue, g, eps);
    stress(P, eps, s);
    F w = msimp_weight(P, zc);
    F sw[6], r[12];
    for (int k = 0; k < 6; ++k) sw[k] = w * s[k];
    divergence(sw, g, v, r);
    for (int row = 0; row < 12; ++row)
      for (int cc = 0; cc < 4; ++cc) {
        const int s2 = block_entry(rowmap, col, conn[4 * e + cc], conn[4 * e + row / 3]);
        atomic_add(&values_t[3 * (size_t)s2 + row % 3], r[row].dx[cc]);
      }
  }
}

// InternalElasticEnergy (src/elliptic/InternalElasticEnergy.hpp:100-158) through
// PhysicsScalarFunction value / gradient_u / gradient_z
// (src/elliptic/PhysicsScalarFunction.hpp:225-283, 365-…, 432-495).
// Any of value/grad_u/grad_z may be NULL.
void orc_energy(int nn, int ne, const double* X, const int* conn, const double* u, const double* z, double E,
                double nu, double p, double rmin, double* value, double* grad_u, double* grad_z) {
  Mesh m{nn, ne, X, conn};
  Params P;
  make_params(E, nu, p, rmin, P);
  auto cfg = workset_config(m);
  if (value) {
    double total = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : total)
    for (int e = 0; e < ne; ++e) {
      double g[12], v, ue[12], zc[4];
      gradient_workset(&cfg[12 * (size_t)e], g, v);
      for (int n = 0; n < 4; ++n) {
        zc[n] = z[conn[4 * e + n]];
        for (int d = 0; d < 3; ++d) ue[3 * n + d] = u[3 * (size_t)conn[4 * e + n] + d];
      }
      double eps[6], s[6];
      strain(ue, g, eps);
      stress(P, eps, s);
      double w = msimp_weight(P, zc);
      double inc = 0.0;
      for (int k = 0; k < 6; ++k) inc += (s[k] * w) * eps[k];
      total += 0.5 * inc * v;
    }
    *value = total;
  }
  if (grad_u) {
    using F = Fad<12>;
    std::memset(grad_u, 0, sizeof(double) * 3 * (size_t)nn);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < ne; ++e) {
      double g[12], v, zc[4];
      gradient_workset(&cfg[12 * (size_t)e], g, v);
      F ue[12];
      for (int n = 0; n < 4; ++n) {
        zc[n] = z[conn[4 * e + n]];
        for (int d = 0; d < 3; ++d) ue[3 * n + d] = F(u[3 * (size_t)conn[4 * e + n] + d], 3 * n + d);
      }
      F eps[6], s[6];
      strain(ue, g, eps);
      stress(P, eps, s);
      double w = msimp_weight(P, zc);
      F inc(0.0);
      for (int k = 0; k < 6; ++k) inc += (s[k] * w) * eps[k];
      F val = inc * (0.5 * v);
      for (int k = 0; k < 12; ++k) atomic_add(&grad_u[3 * (size_t)conn[4 * e + k / 3] + k % 3], val.dx[k]);
    }
  }
  if (grad_z) {
    using F = Fad<4>;
    std::memset(grad_z, 0, sizeof(double) * (size_t)nn);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < ne; ++e) {
      double g[12], v, ue[12];
      gradient_workset(&cfg[12 * (size_t)e], g, v);
      F zc[4];
      for (int n = 0; n < 4; ++n) {
        zc[n] = F(z[conn[4 * e + n]], n);
        for (int d = 0; d < 3; ++d) ue[3 * n + d] = u[3 * (size_t)conn[4 * e + n] + d];
      }
      double eps[6], s[6];
      strain(ue, g, eps);
      stress(P, eps, s);
      F w = msimp_weight(P, zc);
      F inc(0.0);
      for (int k = 0; k < 6; ++k) inc += (w * s[k]) * eps[k];
      F val = inc * (0.5 * v);
      for (int n = 0; n < 4; ++n) atomic_add(&grad_z[conn[4 * e + n]], val.dx[n]);
    }
  }
}

// applyBlockConstraints<3> (src/ApplyConstraints.hpp:11-79), serial.
void orc_apply_constraints(int nn, const int* rowmap, const int* col, double* values, double* rhs, int nbc,
                           const int* dofs, const double* bcvals, double scale) {
  (void)nn;
  for (int b = 0; b < nbc; ++b) {
    const int d = dofs[b];
    const double v = scale * bcvals[b];
    const int rn = d / 3, rd = d % 3;
    for (int s = rowmap[rn]; s < rowmap[rn + 1]; ++s)
      for (int j = 0; j < 3; ++j) {
        const int cn = col[s];
        const size_t ent = 9 * (size_t)s + 3 * rd + j;
        const int cd = 3 * cn + j;
        if (cd == d) {
          values[ent] = 1.0;
        } else {
          rhs[cd] -= values[ent] * v;
          values[ent] = 0.0;
          for (int t = rowmap[cn]; t < rowmap[cn + 1]; ++t)
            if (3 * col[t] + rd == d) values[9 * (size_t)t + 3 * j + rd] = 0.0;
        }
      }
  }
  for (int b = 0; b < nbc; ++b) rhs[dofs[b]] = scale * bcvals[b];
}

static void bcsr_mv(int nn, const int* rowmap, const int* col, const double* A, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int r = 0; r < nn; ++r) {
    double a0 = 0, a1 = 0, a2 = 0;
    for (int s = rowmap[r]; s < rowmap[r + 1]; ++s) {
      const double* b = A + 9 * (size_t)s;
      const double* xc = x + 3 * (size_t)col[s];
      a0 += b[0] * xc[0] + b[1] * xc[1] + b[2] * xc[2];
      a1 += b[3] * xc[0] + b[4] * xc[1] + b[5] * xc[2];
      a2 += b[6] * xc[0] + b[7] * xc[1] + b[8] * xc[2];
    }
    y[3 * (size_t)r] = a0;
    y[3 * (size_t)r + 1] = a1;
    y[3 * (size_t)r + 2] = a2;
  }
}

void orc_matvec(int nn, const int* rowmap, const int* col, const double* A, const double* x, double* y) {
  bcsr_mv(nn, rowmap, col, A, x, y);
}

// Jacobi (3x3 block) preconditioned CG in FP64; returns iterations, *relres = true
// ||b-Ax||/||b||.  x is overwritten (zero start).
int orc_pcg(int nn, const int* rowmap, const int* col, const double* A, const double* b, double* x, double tol,
            int maxit, double* relres) {
  const size_t n = 3 * (size_t)nn;
  std::vector<double> r(b, b + n), zz(n), p(n), q(n), dinv(9 * (size_t)nn);
  for (int row = 0; row < nn; ++row) {
    int s = block_entry(rowmap, col, row, row);
    const double* a = A + 9 * (size_t)s;
    double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
    double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    double* o = &dinv[9 * (size_t)row];
    o[0] = c00 / det; o[1] = (a[2] * a[7] - a[1] * a[8]) / det; o[2] = (a[1] * a[5] - a[2] * a[4]) / det;
    o[3] = c01 / det; o[4] = (a[0] * a[8] - a[2] * a[6]) / det; o[5] = (a[2] * a[3] - a[0] * a[5]) / det;
    o[6] = c02 / det; o[7] = (a[1] * a[6] - a[0] * a[7]) / det; o[8] = (a[0] * a[4] - a[1] * a[3]) / det;
  }
  auto prec = [&](const std::vector<double>& in, std::vector<double>& out) {
#pragma omp parallel for schedule(static)
    for (int row = 0; row < nn; ++row) {
      const double* D = &dinv[9 * (size_t)row];
      const double* v = &in[3 * (size_t)row];
      for (int i = 0; i < 3; ++i) out[3 * (size_t)row + i] = D[3 * i] * v[0] + D[3 * i + 1] * v[1] + D[3 * i + 2] * v[2];
    }
  };
  auto dot = [&](const std::vector<double>& a, const std::vector<double>& c) {
    double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
    for (size_t i = 0; i < n; ++i) s += a[i] * c[i];
    return s;
  };
  std::fill(x, x + n, 0.0);
  std::vector<double> bv(b, b + n);
  const double bb = dot(bv, bv);
  if (bb == 0.0) {
    *relres = 0.0;
    return 0;
  }
  prec(r, zz);
  p = zz;
  double rz = dot(r, zz);
  int it = 0;
  while (it < maxit) {
    if (std::sqrt(dot(r, r) / bb) <= tol) break;
    bcsr_mv(nn, rowmap, col, A, p.data(), q.data());
    const double alpha = rz / dot(p, q);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    prec(r, zz);
    const double rz2 = dot(r, zz);
    const double beta = rz2 / rz;
    rz = rz2;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) p[i] = zz[i] + beta * p[i];
    ++it;
  }
  bcsr_mv(nn, rowmap, col, A, x, q.data());
  double rr = 0.0;
  for (size_t i = 0; i < n; ++i) rr += (b[i] - q[i]) * (b[i] - q[i]);
  *relres = std::sqrt(rr / bb);
  return it;
}

}  // extern "C"

// Stress p-norm criteria, F = ( sum_e vol_e * m_e^p )^(1/p):
//   measure 0: the reference StressPNorm (src/elliptic/StressPNorm.hpp:88-142) with
//     TensorPNorm (src/TensorPNorm.hpp:28-45, 286-318): m_e = |sigma_w|_Voigt, the
//     unit-weight Voigt norm of the SIMP-weighted stress; the functor receives the
//     exponent as an integer (TensorPNormFunctor's `Plato::OrdinalType p`).
//   measure 1: relaxed von Mises, m_e = rho_e^q vm(sigma0_e) with the unpenalized
//     stress sigma0 and vm of VonMisesYieldFunction (src/VonMisesYieldFunction.hpp:53-74).
// Derivatives through forward-mode dual numbers (12 state / 4 control slots) and the
// postEvaluate scaling pow(S,(1-p)/p)/p (TensorNormBase::postEvaluate).
template <typename T>
T vm_of(const T* s) {
  T d01 = s[0] - s[1], d12 = s[1] - s[2], d20 = s[2] - s[0];
  T j2 = 0.5 * (d01 * d01 + d12 * d12 + d20 * d20) + 3.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
  return fad_sqrt(j2);
}

template <typename R, typename S, typename Z>
R pnorm_cell(const Params& P, const S* eps, const Z* zc, double vol, int measure, double pn, double q) {
  S s[6];
  stress(P, eps, s);
  Z rho(0.0);
  for (int n = 0; n < 4; ++n) rho += zc[n];
  rho = rho / 4.0;
  if (measure == 0) {
    // the criterion's own "Penalty Function" (src/Mechanics.hpp:241-276 instantiates
    // StressPNorm per penalty model)
    Z w = P.model != 0 ? penalty_of(P, rho) : P.rmin + (1.0 - P.rmin) * fad_pow(rho, P.p);
    R ss(0.0);
    for (int k = 0; k < 6; ++k) {
      R t = mix<R>(s[k], w);
      ss += t * t;
    }
    return fad_pow(ss, (int)pn / 2.0) * vol;
  }
  Z rq = fad_pow(rho, q);
  R sr[6];
  for (int k = 0; k < 6; ++k) sr[k] = mix<R>(s[k], rq);
  R m = vm_of(sr);
  return fad_pow(m, pn) * vol;
}

extern "C" {

// value; grad_u (3N) and grad_z (N) are the partial derivatives of F (may be NULL)
double orc_pnorm(int nn, int ne, const double* X, const int* conn, const double* u, const double* z, double E,
                 double nu, double p, double rmin, int measure, double pn, double q, double* grad_u,
                 double* grad_z) {
  Mesh m{nn, ne, X, conn};
  Params P;
  make_params(E, nu, p, rmin, P);
  auto cfg = workset_config(m);
  double S = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : S)
  for (int e = 0; e < ne; ++e) {
    double g[12], v, ue[12], zc[4];
    gradient_workset(&cfg[12 * (size_t)e], g, v);
    for (int n = 0; n < 4; ++n) {
      zc[n] = z[conn[4 * e + n]];
      for (int d = 0; d < 3; ++d) ue[3 * n + d] = u[3 * (size_t)conn[4 * e + n] + d];
    }
    double eps[6];
    strain(ue, g, eps);
    S += pnorm_cell<double, double, double>(P, eps, zc, v, measure, pn, q);
  }
  const double pe = measure == 0 ? (double)(int)pn : pn;
  const double scale = std::pow(S, (1.0 - pe) / pe) / pe;
  if (grad_u) {
    using F = Fad<12>;
    std::memset(grad_u, 0, sizeof(double) * 3 * (size_t)nn);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < ne; ++e) {
      double g[12], v, zc[4];
      gradient_workset(&cfg[12 * (size_t)e], g, v);
      F ue[12];
      for (int n = 0; n < 4; ++n) {
        zc[n] = z[conn[4 * e + n]];
        for (int d = 0; d < 3; ++d) ue[3 * n + d] = F(u[3 * (size_t)conn[4 * e + n] + d], 3 * n + d);
      }
      F eps[6];
      strain(ue, g, eps);
      F val = pnorm_cell<F, F, double>(P, eps, zc, v, measure, pn, q);
      for (int k = 0; k < 12; ++k)
        atomic_add(&grad_u[3 * (size_t)conn[4 * e + k / 3] + k % 3], scale * val.dx[k]);
    }
  }
  if (grad_z) {
    using F = Fad<4>;
    std::memset(grad_z, 0, sizeof(double) * (size_t)nn);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < ne; ++e) {
      double g[12], v, ue[12];
      gradient_workset(&cfg[12 * (size_t)e], g, v);
      F zc[4];
      for (int n = 0; n < 4; ++n) {
        zc[n] = F(z[conn[4 * e + n]], n);
        for (int d = 0; d < 3; ++d) ue[3 * n + d] = u[3 * (size_t)conn[4 * e + n] + d];
      }
      double eps[6];
      strain(ue, g, eps);
      F val = pnorm_cell<F, double, F>(P, eps, zc, v, measure, pn, q);
      for (int n = 0; n < 4; ++n) atomic_add(&grad_z[conn[4 * e + n]], scale * val.dx[n]);
    }
  }
  return std::pow(S, 1.0 / pe);
}

double orc_stress_pnorm(int nn, int ne, const double* X, const int* conn, const double* u, const double* z,
                        double E, double nu, double p, double rmin, double pnorm) {
  return orc_pnorm(nn, ne, X, conn, u, z, E, nu, p, rmin, 0, pnorm, 0.0, nullptr, nullptr);
}

// VonMisesYieldFunction<3> (src/VonMisesYieldFunction.hpp:53-74) of n Voigt stresses.
void orc_von_mises(int n, const double* stress6, double* out) {
  for (int c = 0; c < n; ++c) out[c] = vm_of(stress6 + 6 * (size_t)c);
}

// PhysicsScalarFunction::gradient_x of the stress p-norm criteria
// (src/elliptic/PhysicsScalarFunction.hpp:293-355): the cell coordinates seeded as
// SFad<12> through ComputeGradientWorkset, strain, stress, weighting and the cell
// volume (src/elliptic/StressPNorm.hpp:88-142), then the TensorNormBase::postEvaluate
// scaling pow(S,(1-p)/p)/p of the assembled gradient (src/TensorPNorm.hpp:295-305).
double orc_pnorm_gradient_x(int nn, int ne, const double* X, const int* conn, const double* u, const double* z,
                            double E, double nu, double p, double rmin, int measure, double pn, double q,
                            double* gx) {
  using F = Fad<12>;
  Mesh m{nn, ne, X, conn};
  Params P;
  make_params(E, nu, p, rmin, P);
  auto cfgw = workset_config(m);
  std::memset(gx, 0, sizeof(double) * 3 * (size_t)nn);
  double S = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : S)
  for (int e = 0; e < ne; ++e) {
    F cfg[12], g[12], vol;
    double ue[12], zc[4];
    for (int k = 0; k < 12; ++k) cfg[k] = F(cfgw[12 * (size_t)e + k], k);
    for (int n = 0; n < 4; ++n) {
      zc[n] = z[conn[4 * e + n]];
      for (int d = 0; d < 3; ++d) ue[3 * n + d] = u[3 * (size_t)conn[4 * e + n] + d];
    }
    gradient_workset_t(cfg, g, vol);
    double rho = 0.0;
    for (int n = 0; n < 4; ++n) rho += zc[n];
    rho = rho / 4.0;
    F eps[6], sw[6], val;
    if (measure == 0) {
      element_stress_t(P, g, ue, msimp_weight(P, zc), eps, sw);
      F ss(0.0);
      for (int k = 0; k < 6; ++k) ss += sw[k] * sw[k];
      val = fad_pow(ss, (int)pn / 2.0) * vol;
    } else {
      element_stress_t(P, g, ue, std::pow(rho, q), eps, sw);
      val = fad_pow(vm_of(sw), pn) * vol;
    }
    S += val.v;
    for (int k = 0; k < 12; ++k) atomic_add(&gx[3 * (size_t)conn[4 * e + k / 3] + k % 3], val.dx[k]);
  }
  const double pe = measure == 0 ? (double)(int)pn : pn;
  const double scale = std::pow(S, (1.0 - pe) / pe) / pe;
  for (size_t k = 0; k < 3 * (size_t)nn; ++k) gx[k] *= scale;
  return std::pow(S, 1.0 / pe);
}

// (dR/dz)^T lambda per control node: sum_e dw/drho/4 * lambda_e . (vol B^T sigma0(u_e))
// (VectorFunction::gradient_z transposed, applied to a vector: src/PlatoMathHelpers.hpp:58-138)
void orc_dRdz_T_apply(int nn, int ne, const double* X, const int* conn, const double* u, const double* lam,
                      const double* z, double E, double nu, double p, double rmin, double* out) {
  Mesh m{nn, ne, X, conn};
  Params P;
  make_params(E, nu, p, rmin, P);
  auto cfg = workset_config(m);
  std::memset(out, 0, sizeof(double) * (size_t)nn);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < ne; ++e) {
    double g[12], v, ue[12], le[12];
    gradient_workset(&cfg[12 * (size_t)e], g, v);
    Fad<4> zc[4];
    for (int n = 0; n < 4; ++n) {
      zc[n] = Fad<4>(z[conn[4 * e + n]], n);
      for (int d = 0; d < 3; ++d) {
        ue[3 * n + d] = u[3 * (size_t)conn[4 * e + n] + d];
        le[3 * n + d] = lam[3 * (size_t)conn[4 * e + n] + d];
      }
    }
    double eps[6], s[6];
    strain(ue, g, eps);
    stress(P, eps, s);
    Fad<4> w = msimp_weight(P, zc);
    Fad<4> sw[6], r[12];
    for (int k = 0; k < 6; ++k) sw[k] = w * s[k];
    divergence(sw, g, v, r);
    for (int n = 0; n < 4; ++n) {
      double acc = 0.0;
      for (int k = 0; k < 12; ++k) acc += le[k] * r[k].dx[n];
      atomic_add(&out[conn[4 * e + n]], acc);
    }
  }
}

// VectorFunction::gradient_x: dR/dX with the configuration seeded as SFad<12>, stored
// transposed like gradient_z (assemble_transpose_jacobian, src/Assembly.hpp:1422-1485):
// block (config node b, residual node a), entry (k, i) = dR_{a,i}/dX_{b,k}; then the
// boundary pass: the traction load's dependence on the face coordinates (SFad over the
// face vertices; R = K u - f).
void orc_gradient_x(int nn, int ne, const double* X, const int* conn, const double* u, const double* z, double E,
                    double nu, double p, double rmin, int nfaces, const int* face_verts, const double* t,
                    const int* rowmap, const int* col, double* values_t) {
  using F = Fad<12>;
  Mesh m{nn, ne, X, conn};
  Params P;
  make_params(E, nu, p, rmin, P);
  std::memset(values_t, 0, sizeof(double) * 9 * (size_t)rowmap[nn]);
  auto cfgw = workset_config(m);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < ne; ++e) {
    F cfg[12], g[12], vol;
    double ue[12], zc[4];
    for (int k = 0; k < 12; ++k) cfg[k] = F(cfgw[12 * (size_t)e + k], k);
    for (int n = 0; n < 4; ++n) {
      zc[n] = z[conn[4 * e + n]];
      for (int d = 0; d < 3; ++d) ue[3 * n + d] = u[3 * (size_t)conn[4 * e + n] + d];
    }
    gradient_workset_t(cfg, g, vol);
    const double w = msimp_weight(P, zc);
    F eps[6], sw[6], r[12];
    element_stress_t(P, g, ue, w, eps, sw);
    for (int i = 0; i < 3; ++i)
      for (int n = 0; n < 4; ++n) {
        F acc(0.0);
        for (int j = 0; j < 3; ++j) acc += vol * sw[kVoigt[i][j]] * g[3 * n + j];
        r[3 * n + i] = acc;
      }
    for (int row = 0; row < 12; ++row)
      for (int cc = 0; cc < 12; ++cc) {
        const int s2 = block_entry(rowmap, col, conn[4 * e + cc / 3], conn[4 * e + row / 3]);
        atomic_add(&values_t[9 * (size_t)s2 + 3 * (cc % 3) + row % 3], r[row].dx[cc]);
      }
  }
  const double cub = 1.0 / 2.0 / 3.0;
  for (int f = 0; f < nfaces; ++f) {
    using G = Fad<9>;
    const int* v = face_verts + 3 * f;
    G x[3][3];
    for (int a = 0; a < 3; ++a)
      for (int d = 0; d < 3; ++d) x[a][d] = G(X[3 * (size_t)v[a] + d], 3 * a + d);
    G j0[3], j1[3];
    for (int d = 0; d < 3; ++d) {
      j0[d] = x[0][d] - x[2][d];
      j1[d] = x[1][d] - x[2][d];
    }
    G c23 = j0[1] * j1[2] - j0[2] * j1[1];
    G c31 = j0[2] * j1[0] - j0[0] * j1[2];
    G c12 = j0[0] * j1[1] - j0[1] * j1[0];
    G wf = fad_sqrt(c23 * c23 + c31 * c31 + c12 * c12) * (-1.0 * cub);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        const int s2 = block_entry(rowmap, col, v[b], v[a]);
        for (int k = 0; k < 3; ++k)
          for (int i = 0; i < 3; ++i) values_t[9 * (size_t)s2 + 3 * k + i] += wf.dx[3 * b + k] * t[i];
      }
  }
}

// PhysicsScalarFunction::gradient_x of the internal elastic energy: d/dX of
// sum_e 0.5 vol sigma_w . eps with the configuration seeded as SFad<12>; g (3N).
void orc_energy_gradient_x(int nn, int ne, const double* X, const int* conn, const double* u, const double* z,
                           double E, double nu, double p, double rmin, double* gx) {
  using F = Fad<12>;
  Mesh m{nn, ne, X, conn};
  Params P;
  make_params(E, nu, p, rmin, P);
  std::memset(gx, 0, sizeof(double) * 3 * (size_t)nn);
  auto cfgw = workset_config(m);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < ne; ++e) {
    F cfg[12], g[12], vol;
    double ue[12], zc[4];
    for (int k = 0; k < 12; ++k) cfg[k] = F(cfgw[12 * (size_t)e + k], k);
    for (int n = 0; n < 4; ++n) {
      zc[n] = z[conn[4 * e + n]];
      for (int d = 0; d < 3; ++d) ue[3 * n + d] = u[3 * (size_t)conn[4 * e + n] + d];
    }
    gradient_workset_t(cfg, g, vol);
    const double w = msimp_weight(P, zc);
    F eps[6], sw[6], inc(0.0);
    element_stress_t(P, g, ue, w, eps, sw);
    for (int k = 0; k < 6; ++k) inc += sw[k] * eps[k];
    F val = inc * vol * 0.5;
    for (int k = 0; k < 12; ++k) atomic_add(&gx[3 * (size_t)conn[4 * e + k / 3] + k % 3], val.dx[k]);
  }
}

// Plato::Geometric::Volume (src/geometric/Volume.hpp:49-96) through
// GeometryScalarFunction: V = sum_e (|det J|/6) w(mean z_e) with the penalty
// (model, p, rmin, len); dV/dz by SFad<4> on the controls, dV/dX by SFad<12> on the
// cell coordinates.  dvdz (N) and dvdx (3N) may be NULL.
double orc_volume(int nn, int ne, const double* X, const int* conn, const double* z, int model, double p,
                  double rmin, double len, double* dvdz, double* dvdx) {
  Mesh m{nn, ne, X, conn};
  Params P;
  make_params(1.0, 0.3, p, rmin, P);
  P.model = model;
  P.len = len;
  auto cfgw = workset_config(m);
  double total = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : total)
  for (int e = 0; e < ne; ++e) {
    double g[12], v, zc[4];
    gradient_workset(&cfgw[12 * (size_t)e], g, v);
    for (int n = 0; n < 4; ++n) zc[n] = z[conn[4 * e + n]];
    total += v * msimp_weight(P, zc);
  }
  if (dvdz) {
    std::memset(dvdz, 0, sizeof(double) * (size_t)nn);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < ne; ++e) {
      double g[12], v;
      gradient_workset(&cfgw[12 * (size_t)e], g, v);
      Fad<4> zc[4];
      for (int n = 0; n < 4; ++n) zc[n] = Fad<4>(z[conn[4 * e + n]], n);
      Fad<4> val = msimp_weight(P, zc) * v;
      for (int n = 0; n < 4; ++n) atomic_add(&dvdz[conn[4 * e + n]], val.dx[n]);
    }
  }
  if (dvdx) {
    std::memset(dvdx, 0, sizeof(double) * 3 * (size_t)nn);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < ne; ++e) {
      Fad<12> cfg[12], g[12], v;
      double zc[4];
      for (int k = 0; k < 12; ++k) cfg[k] = Fad<12>(cfgw[12 * (size_t)e + k], k);
      for (int n = 0; n < 4; ++n) zc[n] = z[conn[4 * e + n]];
      gradient_workset_t(cfg, g, v);
      Fad<12> val = v * msimp_weight(P, zc);
      for (int k = 0; k < 12; ++k) atomic_add(&dvdx[3 * (size_t)conn[4 * e + k / 3] + k % 3], val.dx[k]);
    }
  }
  return total;
}

}  // extern "C"
