"""Drop-in surface of the reference's elliptic managers over torch device tensors.

Mirrors (same names, argument meaning and error behaviour -- failures raise):
  Plato::Elliptic::VectorFunction         src/elliptic/VectorFunction.hpp:231-689
  Plato::Elliptic::PhysicsScalarFunction  src/elliptic/PhysicsScalarFunction.hpp:191-495
  Plato::Elliptic::Problem                src/elliptic/Problem.hpp:259-468
Every compute call goes through libplato_mi355x.so; torch provides device memory and
the stream.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _capi
from .mesh import TetMesh


def _dptr(t: torch.Tensor):
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return C.c_void_p(t.data_ptr())


def _dvec(t, n: int, what: str):
    """Device pointer of a contiguous float64 vector of exactly n entries; ValueError
    otherwise (the library reads or writes n doubles through the pointer)."""
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{what}: expected a torch tensor, got {type(t).__name__}")
    if t.dtype != torch.float64:
        raise ValueError(f"{what}: expected float64, got {t.dtype}")
    if t.numel() != n:
        raise ValueError(f"{what}: expected {n} entries, got {t.numel()}")
    return _dptr(t)


def _hptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


PENALTY_MODELS = {"SIMP": 0, "RAMP": 1, "Heaviside": 2, "NoPenalty": 3}


def penalty_model_id(name: str) -> int:
    """Index of a "Penalty Function" Type (src/Simp.hpp, src/Ramp.hpp, src/Heaviside.hpp,
    src/NoPenalty.hpp)."""
    if name not in PENALTY_MODELS:
        raise ValueError(f"penalty model must be one of {sorted(PENALTY_MODELS)}")
    return PENALTY_MODELS[name]


@dataclass
class CrsMatrix:
    """Block CRS (src/alg/CrsMatrix.hpp:20-111): rowMap, columnIndices, entries."""
    row_map: torch.Tensor
    column_indices: torch.Tensor
    entries: torch.Tensor
    block_rows: int
    block_cols: int

    def rowMap(self):  # noqa: N802 - reference name
        return self.row_map

    def columnIndices(self):  # noqa: N802
        return self.column_indices

    def entries_(self):
        return self.entries


@dataclass
class EssentialBC:
    """One "Essential Boundary Conditions" entry: node set, dof index, value."""
    nodes: np.ndarray
    dof: int
    value: float = 0.0


@dataclass
class Traction:
    """One "Natural Boundary Conditions" Uniform traction on a side set."""
    face_verts: np.ndarray
    face_elems: np.ndarray
    traction: tuple


@dataclass
class ElasticModel:
    """Handle owner: mesh + isotropic material + MSIMP penalty + BCs + device graph."""
    mesh: TetMesh
    youngs_modulus: float = 1.0
    poissons_ratio: float = 0.3
    penalty_exponent: float = 3.0
    penalty_minimum: float = 1e-9
    essential_bcs: list = field(default_factory=list)
    tractions: list = field(default_factory=list)
    graph: tuple | None = None  # optional (rowmap, colidx) host arrays
    device: str = "cuda"
    box: tuple | None = None  # (nx, ny, nz) hexes of a structured box mesh: enables multigrid
    penalty_model: str = "SIMP"  # "Penalty Function" Type: SIMP, RAMP, Heaviside, NoPenalty
    regularization_length: float = 1.0  # Heaviside only
    box_optional: bool = False  # a box the library refuses (nodes off the lattice) leaves box = None

    def __post_init__(self):
        h = C.c_void_p()
        coords = np.ascontiguousarray(self.mesh.coords, np.float64)
        conn = np.ascontiguousarray(self.mesh.conn, np.int32)
        _capi.call("pa_create", C.byref(h), self.mesh.num_nodes, self.mesh.num_elems, _hptr(coords), _hptr(conn))
        self._h = h
        _capi.call("pa_set_material", h, float(self.youngs_modulus), float(self.poissons_ratio))
        _capi.call("pa_set_penalty_model", h, penalty_model_id(self.penalty_model), float(self.penalty_exponent),
                   float(self.penalty_minimum), float(self.regularization_length))
        for bc in self.essential_bcs:
            nodes = np.ascontiguousarray(bc.nodes, np.int32)
            _capi.call("pa_add_essential_bc", h, len(nodes), _hptr(nodes), int(bc.dof), float(bc.value))
        for tr in self.tractions:
            fv = np.ascontiguousarray(tr.face_verts, np.int32)
            fe = np.ascontiguousarray(tr.face_elems, np.int32)
            t = np.ascontiguousarray(tr.traction, np.float64)
            _capi.call("pa_add_traction", h, len(fe), _hptr(fv), _hptr(fe), _hptr(t))
        if self.graph is not None:
            rm = np.ascontiguousarray(self.graph[0], np.int32)
            ci = np.ascontiguousarray(self.graph[1], np.int32)
            _capi.call("pa_set_graph", h, _hptr(rm), _hptr(ci))
        _capi.call("pa_finalize", h)
        if self.box is not None:
            try:
                _capi.call("pa_set_box", h, *(int(v) for v in self.box))
            except _capi.PaError as e:
                # the finalized handle stays usable without a box (general assembly)
                if not (self.box_optional and e.status == _capi.PA_ERR_MESH):
                    raise
                self.box = None
        nr, nb = C.c_int(), C.c_int()
        _capi.call("pa_graph_size", h, C.byref(nr), C.byref(nb))
        self.num_rows, self.nnz_blocks = nr.value, nb.value
        self.row_map = torch.empty(self.num_rows + 1, dtype=torch.int32, device=self.device)
        self.column_indices = torch.empty(self.nnz_blocks, dtype=torch.int32, device=self.device)
        self._stream()
        _capi.call("pa_copy_graph", h, _dptr(self.row_map), _dptr(self.column_indices), 1)

    @property
    def handle(self):
        return self._h

    @property
    def box_assembly(self) -> bool:
        """True when the handle assembles the Jacobian on the lattice path (pa_box_assembly)."""
        v = C.c_int(0)
        _capi.call("pa_box_assembly", self._h, C.byref(v))
        return bool(v.value)

    @property
    def num_nodes(self) -> int:
        return self.mesh.num_nodes

    @property
    def num_dofs(self) -> int:
        return 3 * self.mesh.num_nodes

    def _stream(self):
        _capi.call("pa_set_stream", self._h, C.c_void_p(torch.cuda.current_stream().cuda_stream))

    def _uz(self, u, z):
        """Checked device pointers of a state (3N float64) and a control (N float64)."""
        return _dvec(u, self.num_dofs, "state u"), _dvec(z, self.num_nodes, "control z")

    def close(self):
        if getattr(self, "_h", None):
            _capi.lib().pa_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class VectorFunction:
    """Plato::Elliptic::VectorFunction<Mechanics<3>> for the elastostatics residual."""

    def __init__(self, model: ElasticModel):
        self.m = model

    def _check(self, u, z):
        if u is not None:
            _dvec(u, self.m.num_dofs, "state u")
        _dvec(z, self.m.num_nodes, "control z")

    def value(self, u: torch.Tensor, z: torch.Tensor, time_step: float = 0.0) -> torch.Tensor:
        self._check(u, z)
        R = torch.empty(self.m.num_dofs, dtype=torch.float64, device=u.device)
        self.m._stream()
        _capi.call("pa_vf_value", self.m.handle, _dptr(u), _dptr(z), _dptr(R))
        return R

    def gradient_u(self, u: torch.Tensor, z: torch.Tensor, time_step: float = 0.0,
                   residual_out: torch.Tensor | None = None) -> CrsMatrix:
        self._check(u, z)
        vals = torch.empty(9 * self.m.nnz_blocks, dtype=torch.float64, device=z.device)
        self.m._stream()
        _capi.call("pa_vf_gradient_u", self.m.handle, _dptr(u) if u is not None else None, _dptr(z), _dptr(vals),
                   _dvec(residual_out, self.m.num_dofs, "residual_out") if residual_out is not None else None)
        return CrsMatrix(self.m.row_map, self.m.column_indices, vals, 3, 3)

    gradient_u_T = gradient_u  # the elastostatic Jacobian is symmetric

    def gradient_z(self, u: torch.Tensor, z: torch.Tensor, time_step: float = 0.0) -> CrsMatrix:
        self._check(u, z)
        vals = torch.empty(3 * self.m.nnz_blocks, dtype=torch.float64, device=z.device)
        self.m._stream()
        _capi.call("pa_vf_gradient_z", self.m.handle, _dptr(u), _dptr(z), _dptr(vals))
        return CrsMatrix(self.m.row_map, self.m.column_indices, vals, 1, 3)

    def gradient_x(self, u: torch.Tensor, z: torch.Tensor, time_step: float = 0.0) -> CrsMatrix:
        """dR/dX, transposed 3x3 blocks: block (config node, residual node)."""
        self._check(u, z)
        vals = torch.empty(9 * self.m.nnz_blocks, dtype=torch.float64, device=z.device)
        self.m._stream()
        _capi.call("pa_vf_gradient_x", self.m.handle, _dptr(u), _dptr(z), _dptr(vals))
        return CrsMatrix(self.m.row_map, self.m.column_indices, vals, 3, 3)


class PhysicsScalarFunction:
    """Plato::Elliptic::PhysicsScalarFunction for "Internal Elastic Energy"."""

    def __init__(self, model: ElasticModel, function_type: str = "Internal Elastic Energy"):
        if function_type != "Internal Elastic Energy":
            raise ValueError(f"unsupported scalar function type {function_type!r}")
        self.m = model

    def value(self, u: torch.Tensor, z: torch.Tensor, time_step: float = 0.0) -> float:
        v = C.c_double()
        up, zp = self.m._uz(u, z)
        self.m._stream()
        _capi.call("pa_sf_value", self.m.handle, up, zp, C.byref(v))
        return v.value

    def gradient_u(self, u: torch.Tensor, z: torch.Tensor, step: int = 0, time_step: float = 0.0) -> torch.Tensor:
        g = torch.empty(self.m.num_dofs, dtype=torch.float64, device=u.device)
        up, zp = self.m._uz(u, z)
        self.m._stream()
        _capi.call("pa_sf_gradient_u", self.m.handle, up, zp, _dptr(g))
        return g

    def gradient_z(self, u: torch.Tensor, z: torch.Tensor, time_step: float = 0.0) -> torch.Tensor:
        g = torch.empty(self.m.num_nodes, dtype=torch.float64, device=u.device)
        up, zp = self.m._uz(u, z)
        self.m._stream()
        _capi.call("pa_sf_gradient_z", self.m.handle, up, zp, _dptr(g), None)
        return g

    def gradient_x(self, u: torch.Tensor, z: torch.Tensor, time_step: float = 0.0) -> torch.Tensor:
        """Partial derivative of the energy in the nodal coordinates (3N)."""
        g = torch.empty(self.m.num_dofs, dtype=torch.float64, device=u.device)
        up, zp = self.m._uz(u, z)
        self.m._stream()
        _capi.call("pa_sf_gradient_x", self.m.handle, up, zp, _dptr(g))
        return g


class Volume:
    """Plato::Geometric::Volume (src/geometric/Volume.hpp:49-96) with its own "Penalty
    Function" (default NoPenalty): V = sum_e vol_e w(mean z_e); gradient_z, gradient_x
    as GeometryScalarFunction (src/geometric/GeometryScalarFunction.cpp)."""

    def __init__(self, model: ElasticModel, penalty_model: str = "NoPenalty", exponent: float = 1.0,
                 minimum_value: float = 0.0, regularization_length: float = 1.0):
        self.m = model
        self.args = (penalty_model_id(penalty_model), float(exponent), float(minimum_value),
                     float(regularization_length))

    def _eval(self, z: torch.Tensor, want_z: bool, want_x: bool):
        v = C.c_double()
        zp = _dvec(z, self.m.num_nodes, "control z")
        gz = torch.empty(self.m.num_nodes, dtype=torch.float64, device=z.device) if want_z else None
        gx = torch.empty(self.m.num_dofs, dtype=torch.float64, device=z.device) if want_x else None
        self.m._stream()
        _capi.call("pa_volume", self.m.handle, zp, *self.args, C.byref(v),
                   _dptr(gz) if gz is not None else None, _dptr(gx) if gx is not None else None)
        return v.value, gz, gx

    def value(self, z: torch.Tensor) -> float:
        return self._eval(z, False, False)[0]

    def gradient_z(self, z: torch.Tensor) -> torch.Tensor:
        return self._eval(z, True, False)[1]

    def gradient_x(self, z: torch.Tensor) -> torch.Tensor:
        return self._eval(z, False, True)[2]


class Problem:
    """Plato::Elliptic::Problem (self-adjoint compliance)."""

    def __init__(self, model: ElasticModel, tolerance: float = 1e-10, max_iters: int = 200000,
                 check_every: int = 16, precision: str = "fp64", inner_tolerance: float = 0.1,
                 preconditioner: str = "jacobi"):
        """precision "fp64": assembled block-CRS PCG in FP64.  "mixed": PCG with FP64
        vectors and an FP32 matrix-free operator, its residual replaced by the true FP64
        residual f - K u each time it has dropped by `inner_tolerance`.
        preconditioner "jacobi": 3x3 block Jacobi; "multigrid": geometric V-cycle on the
        model's box (ElasticModel(box=(nx, ny, nz)))."""
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        if preconditioner not in ("jacobi", "multigrid"):
            raise ValueError("preconditioner must be 'jacobi' or 'multigrid'")
        self.m = model
        self.opts = _capi.SolverOpts(tolerance, max_iters, check_every, 0 if precision == "fp64" else 1, 0,
                                     inner_tolerance, 0 if preconditioner == "jacobi" else 1)
        self.stats = _capi.SolverStats()
        self._vals = None
        self.state = None

    def solution(self, z: torch.Tensor, u0: torch.Tensor | None = None) -> torch.Tensor:
        """Problem::solution; `u0` warm-starts the solve (any precision and preconditioner;
        its constrained dofs are reset to their prescribed values)."""
        zp = _dvec(z, self.m.num_nodes, "control z")
        if u0 is not None:
            if u0.numel() != self.m.num_dofs:
                raise ValueError(f"u0: expected {self.m.num_dofs} entries, got {u0.numel()}")
            u = u0.detach().to(device=z.device, dtype=torch.float64).clone().contiguous()
            self.opts.use_x0 = 1
        else:
            u = torch.empty(self.m.num_dofs, dtype=torch.float64, device=z.device)
            self.opts.use_x0 = 0
        vals = None
        if self.opts.precision == 0:
            if self._vals is None:
                self._vals = torch.empty(9 * self.m.nnz_blocks, dtype=torch.float64, device=z.device)
            vals = _dptr(self._vals)
        self.m._stream()
        _capi.call("pa_solution", self.m.handle, zp, _dptr(u), vals, C.byref(self.opts), C.byref(self.stats))
        self.state = u
        return u

    def solution_deflated(self, z: torch.Tensor, W: torch.Tensor, u0: torch.Tensor | None = None) -> torch.Tensor:
        """Problem::solution by FP64 block-Jacobi PCG deflated by the rows of W (k x 3N,
        k <= 16; pa_solution_deflated): the start is corrected by the Galerkin solution
        in span(W) and the search directions stay A-orthogonal to it."""
        zp = _dvec(z, self.m.num_nodes, "control z")
        W = W.detach().to(device=z.device, dtype=torch.float64).contiguous()
        if W.dim() != 2 or W.shape[1] != self.m.num_dofs:
            raise ValueError(f"W: expected (k, {self.m.num_dofs}), got {tuple(W.shape)}")
        if u0 is not None:
            if u0.numel() != self.m.num_dofs:
                raise ValueError(f"u0: expected {self.m.num_dofs} entries, got {u0.numel()}")
            u = u0.detach().to(device=z.device, dtype=torch.float64).clone().contiguous()
        else:
            u = torch.empty(self.m.num_dofs, dtype=torch.float64, device=z.device)
        opts = _capi.SolverOpts(self.opts.tolerance, self.opts.max_iters, self.opts.check_every, 0,
                                1 if u0 is not None else 0, self.opts.inner_tolerance, 0)
        if self._vals is None:
            self._vals = torch.empty(9 * self.m.nnz_blocks, dtype=torch.float64, device=z.device)
        self.m._stream()
        _capi.call("pa_solution_deflated", self.m.handle, zp, _dptr(u), _dptr(self._vals), _dptr(W), W.shape[0],
                   C.byref(opts), C.byref(self.stats))
        self.state = u
        return u

    def criterionValue(self, z: torch.Tensor) -> float:  # noqa: N802
        u = self.solution(z)
        return PhysicsScalarFunction(self.m).value(u, z)

    def criterionGradient(self, z: torch.Tensor, u: torch.Tensor | None = None):  # noqa: N802
        if u is None:
            u = self.solution(z)
        g = torch.empty(self.m.num_nodes, dtype=torch.float64, device=z.device)
        v = C.c_double()
        up, zp = self.m._uz(u, z)
        self.m._stream()
        _capi.call("pa_compliance_gradient", self.m.handle, up, zp, _dptr(g), C.byref(v))
        return g, v.value


def solve(model: ElasticModel, values: torch.Tensor, b: torch.Tensor, tolerance: float = 1e-10,
          max_iters: int = 200000, x0: torch.Tensor | None = None):
    """AbstractSolver::solve(A, x, b) on a constrained matrix of the model's graph."""
    vp = _dvec(values, 9 * model.nnz_blocks, "values")
    bp = _dvec(b, model.num_dofs, "b")
    if x0 is not None:
        _dvec(x0, model.num_dofs, "x0")
    x = x0.clone() if x0 is not None else torch.empty_like(b)
    opts = _capi.SolverOpts(tolerance, max_iters, 16, 0, 1 if x0 is not None else 0, 1e-4, 0)
    st = _capi.SolverStats()
    model._stream()
    _capi.call("pa_solve", model.handle, vp, _dptr(x), bp, C.byref(opts), C.byref(st))
    return x, st


def apply_constraints(model: ElasticModel, values: torch.Tensor, rhs: torch.Tensor, scale: float = 1.0):
    vp, rp = _dvec(values, 9 * model.nnz_blocks, "values"), _dvec(rhs, model.num_dofs, "rhs")
    model._stream()
    _capi.call("pa_apply_constraints", model.handle, vp, rp, float(scale))
