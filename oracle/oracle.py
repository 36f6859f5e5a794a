"""ORACLE -- test infrastructure only (tests/, __graft_entry__.smoke(), bench.py's
cpu_baseline).  ctypes wrapper of oracle/_build/liboracle.so, the CPU restatement of
the reference's Tet4 elastostatics path (see oracle.cpp for the function-by-function
citations and what pins each function).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "_build" / "liboracle.so"
_lib = None
# (instruction-set level, /proc/cpuinfo flags it needs), best first
_ISA_FLAGS = (("x86-64-v4", {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"}),
              ("x86-64-v3", {"avx2", "fma", "bmi2", "movbe"}))


def select_isa() -> str:
    """Before the first call: load the oracle build for the best instruction-set level
    this host supports (for CPU timing); returns its name.  The default build is
    x86-64-v2."""
    global LIB
    if _lib is not None:
        raise RuntimeError("oracle already loaded")
    try:
        flags = set()
        for line in open("/proc/cpuinfo"):
            if line.startswith("flags"):
                flags = set(line.split(":", 1)[1].split())
                break
    except OSError:
        flags = set()
    for isa, need in _ISA_FLAGS:
        cand = LIB.parent / f"liboracle_{isa}.so"
        if need <= flags and cand.exists():
            LIB = cand
            return isa
    return "x86-64-v2"


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise ImportError(f"{LIB} not built (python -m platoanalyze_amd.build)")
        _lib = C.CDLL(str(LIB))
        _lib.orc_stress_pnorm.restype = C.c_double
        _lib.orc_pnorm.restype = C.c_double
        _lib.orc_volume.restype = C.c_double
        _lib.orc_pnorm_gradient_x.restype = C.c_double
    return _lib


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _f64(a):
    return np.ascontiguousarray(a, np.float64)


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def num_threads() -> int:
    return int(lib().orc_num_threads())


def graph(conn, nn):
    conn = _i32(conn)
    ne = conn.shape[0]
    rm = np.empty(nn + 1, np.int32)
    nnz = C.c_int()
    lib().orc_graph(nn, ne, _p(conn), _p(rm), None, C.byref(nnz))
    ci = np.empty(nnz.value, np.int32)
    lib().orc_graph(nn, ne, _p(conn), _p(rm), _p(ci), C.byref(nnz))
    return rm, ci


class Model:
    """Mesh + material + penalty (+ traction faces) for the oracle calls."""

    def __init__(self, coords, conn, E=1.0, nu=0.3, p=3.0, rmin=1e-9, face_verts=None, traction=None,
                 penalty_model=0, length=1.0):
        self.X = _f64(coords).reshape(-1, 3)
        self.conn = _i32(conn).reshape(-1, 4)
        self.nn, self.ne = self.X.shape[0], self.conn.shape[0]
        self.E, self.nu, self.p, self.rmin = float(E), float(nu), float(p), float(rmin)
        self.fv = None if face_verts is None else _i32(face_verts).reshape(-1, 3)
        self.t = None if traction is None else _f64(traction)
        self.model, self.len = int(penalty_model), float(length)

    def _m(self):
        return (self.nn, self.ne, _p(self.X), _p(self.conn))

    def _mat(self):
        lib().orc_set_penalty_model(self.model, C.c_double(self.len))  # penalty of this model's calls
        return (C.c_double(self.E), C.c_double(self.nu), C.c_double(self.p), C.c_double(self.rmin))

    def volume(self, z, model=3, p=1.0, rmin=0.0, length=1.0):
        """Geometric volume sum_e vol_e w(mean z_e) and its gradients (dV/dz, dV/dX)."""
        gz, gx = np.empty(self.nn), np.empty(3 * self.nn)
        v = lib().orc_volume(*self._m(), _p(_f64(z)), int(model), C.c_double(p), C.c_double(rmin),
                             C.c_double(length), _p(gz), _p(gx))
        return float(v), gz, gx

    def residual(self, u, z):
        R = np.empty(3 * self.nn)
        nf = 0 if self.fv is None else self.fv.shape[0]
        lib().orc_residual(*self._m(), _p(_f64(u)), _p(_f64(z)), *self._mat(), nf, _p(self.fv), _p(self.t), _p(R))
        return R

    def jacobian(self, u, z, rowmap, colidx):
        v = np.empty(9 * int(rowmap[-1]))
        lib().orc_jacobian(*self._m(), _p(None if u is None else _f64(u)), _p(_f64(z)), *self._mat(),
                           _p(_i32(rowmap)), _p(_i32(colidx)), _p(v))
        return v

    def gradient_z(self, u, z, rowmap, colidx):
        v = np.empty(3 * int(rowmap[-1]))
        lib().orc_gradient_z(*self._m(), _p(_f64(u)), _p(_f64(z)), *self._mat(), _p(_i32(rowmap)),
                             _p(_i32(colidx)), _p(v))
        return v

    def energy(self, u, z, want_grad_u=True, want_grad_z=True):
        val = C.c_double()
        gu = np.empty(3 * self.nn) if want_grad_u else None
        gz = np.empty(self.nn) if want_grad_z else None
        lib().orc_energy(*self._m(), _p(_f64(u)), _p(_f64(z)), *self._mat(), C.byref(val), _p(gu), _p(gz))
        return val.value, gu, gz

    def gradient_x(self, u, z, rowmap, colidx):
        """dR/dX, transposed 3x3 blocks: block (config node b, residual node a), entry
        (k, i) = dR_{a,i}/dX_{b,k} (includes the traction's dependence on the faces)."""
        v = np.empty(9 * int(rowmap[-1]))
        nf = 0 if self.fv is None else self.fv.shape[0]
        lib().orc_gradient_x(*self._m(), _p(_f64(u)), _p(_f64(z)), *self._mat(), nf, _p(self.fv), _p(self.t),
                             _p(_i32(rowmap)), _p(_i32(colidx)), _p(v))
        return v

    def energy_gradient_x(self, u, z):
        """d/dX of the internal elastic energy at fixed u and z (3N)."""
        g = np.empty(3 * self.nn)
        lib().orc_energy_gradient_x(*self._m(), _p(_f64(u)), _p(_f64(z)), *self._mat(), _p(g))
        return g

    def cell_worksets(self, u, ncells):
        g = np.empty((ncells, 4, 3))
        vol = np.empty(ncells)
        eps = np.empty((ncells, 6))
        sig = np.empty((ncells, 6))
        res = np.empty((ncells, 12))
        lib().orc_cell_worksets(*self._m(), _p(_f64(u)), C.c_double(self.E), C.c_double(self.nu), ncells,
                                _p(g), _p(vol), _p(eps), _p(sig), _p(res))
        return g, vol, eps, sig, res

    def pnorm(self, u, z, measure, pn, q=0.0, want_grad_u=True, want_grad_z=True):
        """F = (sum vol m^p)^(1/p); measure 0 = reference StressPNorm (Voigt norm of the
        SIMP-weighted stress), 1 = relaxed von Mises rho^q vm(sigma0).  Returns F and the
        partial derivatives dF/du, dF/dz (None when not requested)."""
        gu = np.empty(3 * self.nn) if want_grad_u else None
        gz = np.empty(self.nn) if want_grad_z else None
        v = lib().orc_pnorm(*self._m(), _p(_f64(u)), _p(_f64(z)), *self._mat(), int(measure), C.c_double(pn),
                            C.c_double(q), _p(gu), _p(gz))
        return float(v), gu, gz

    def pnorm_gradient_x(self, u, z, measure, pn, q=0.0):
        """dF/dX of the stress p-norm at fixed u and z (3N); returns (F, dF/dX)."""
        g = np.empty(3 * self.nn)
        v = lib().orc_pnorm_gradient_x(*self._m(), _p(_f64(u)), _p(_f64(z)), *self._mat(), int(measure),
                                       C.c_double(pn), C.c_double(q), _p(g))
        return float(v), g

    def dRdz_T_apply(self, u, lam, z):
        out = np.empty(self.nn)
        lib().orc_dRdz_T_apply(*self._m(), _p(_f64(u)), _p(_f64(lam)), _p(_f64(z)), *self._mat(), _p(out))
        return out

    def stress_pnorm(self, u, z, pnorm):
        return float(lib().orc_stress_pnorm(*self._m(), _p(_f64(u)), _p(_f64(z)), *self._mat(),
                                            C.c_double(pnorm)))


def von_mises(stress6):
    """VonMisesYieldFunction<3> of Voigt stresses [n][6]."""
    s = _f64(stress6).reshape(-1, 6)
    out = np.empty(s.shape[0])
    lib().orc_von_mises(s.shape[0], _p(s), _p(out))
    return out


def apply_constraints(rowmap, colidx, values, rhs, dofs, bcvals, scale=1.0):
    values = _f64(values).copy()
    rhs = _f64(rhs).copy()
    dofs = _i32(dofs)
    bcvals = _f64(bcvals)
    lib().orc_apply_constraints(len(rowmap) - 1, _p(_i32(rowmap)), _p(_i32(colidx)), _p(values), _p(rhs),
                                len(dofs), _p(dofs), _p(bcvals), C.c_double(scale))
    return values, rhs


def matvec(rowmap, colidx, values, x):
    y = np.empty(3 * (len(rowmap) - 1))
    lib().orc_matvec(len(rowmap) - 1, _p(_i32(rowmap)), _p(_i32(colidx)), _p(_f64(values)), _p(_f64(x)), _p(y))
    return y


def pcg(rowmap, colidx, values, b, tol=1e-10, maxit=100000):
    x = np.empty(3 * (len(rowmap) - 1))
    rel = C.c_double()
    it = lib().orc_pcg(len(rowmap) - 1, _p(_i32(rowmap)), _p(_i32(colidx)), _p(_f64(values)), _p(_f64(b)),
                       _p(x), C.c_double(tol), int(maxit), C.byref(rel))
    return x, int(it), rel.value


def to_dense(rowmap, colidx, values, bs=3):
    """Dense matrix of a bs x bs block CRS (PlatoTestHelpers toFull analogue)."""
    n = len(rowmap) - 1
    A = np.zeros((bs * n, bs * n))
    v = np.asarray(values).reshape(-1, bs, bs)
    for r in range(n):
        for s in range(rowmap[r], rowmap[r + 1]):
            c = colidx[s]
            A[bs * r:bs * r + bs, bs * c:bs * c + bs] = v[s]
    return A
