"""Input decks: Teuchos XML parameter lists and the elliptic mechanics problem they define.

`parse_parameter_list` reads the XML form the reference's decks and unit tests use
(`<ParameterList name=...>` / `<Parameter name=... type=... value=...>`, as consumed by
Teuchos::getParametersFromXmlString) into nested `ParameterList` dicts with typed values.

`ElasticProblem` is Plato::Elliptic::Problem<Mechanics<3>> built from such a deck and an
Exodus mesh (src/PlatoProblemFactory.hpp:70-83, src/elliptic/Problem.hpp:95-130, 620-700):

* "Physics" Mechanical, "PDE Constraint" Elliptic, "Self-Adjoint" (default false);
* "Spatial Model" / "Domains": every domain's "Element Block" must be a block of the
  mesh and all domains name one "Material Model" (one material per handle);
* "Material Models" / <name> / "Isotropic Linear Elastic": "Youngs Modulus",
  "Poissons Ratio" (src/ElasticModelFactory.hpp:43-79);
* "Elliptic" / "Penalty Function": "Type" SIMP (default), RAMP, Heaviside, NoPenalty with
  "Exponent", "Minimum Value", "Regularization Length" and the reference defaults
  (src/Simp.hpp:41-52, src/Ramp.hpp:38-42, src/Heaviside.hpp:46-51);
* "Essential Boundary Conditions": sublists of "Type" Zero Value / Fixed Value, "Index",
  "Sides" (a node set), "Value" (src/EssentialBCs.hpp:225-252);
* "Natural Boundary Conditions": sublists of "Type" Uniform, "Sides" (a side set) and
  "Values" (3 components) or "Index" + "Value" (src/NaturalBCs.hpp:110-210);
* "Criteria": sublists of "Type" Scalar Function with "Scalar Function Type" Internal
  Elastic Energy or Stress P-Norm ("Exponent"), each with its own "Penalty Function"
  (src/Mechanics.hpp:195-215, 400-420); Volume with "Linear" true
  (src/Geometrical.hpp:25-45);
* "Linear Solver": "Iterations", "Tolerance"; the Trilinos/AmgX keys ("Solver Stack",
  "Solver Package", ...) select libraries this build replaces with its own PCG and are
  ignored; "Precision" (FP64 | Mixed) and "Preconditioner" (Jacobi | Multigrid) choose
  among this library's solvers.

A criterion whose penalty differs from the residual's is evaluated on a second handle of
the same mesh carrying that penalty, while its adjoint solve and dR/dz, dR/dX products
use the residual's handle, as the reference evaluates each criterion with its own
penalty model against one PDE.
"""
from __future__ import annotations

import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np
import torch

from .elliptic import ElasticModel, EssentialBC, PhysicsScalarFunction, Problem, Traction, Volume, VectorFunction
from .elliptic import apply_constraints, solve
from .exodus import ExodusMesh
from .mesh import infer_box  # noqa: F401  (part of this module's interface)


_MISSING = object()


class ParameterList(dict):
    """Ordered name -> value map; sublists are ParameterList values."""

    def __init__(self, name: str = ""):
        super().__init__()
        self.name = name

    def get(self, key: str, default=_MISSING):  # Teuchos get: a missing key without default is an error
        if key in self:
            return self[key]
        if default is _MISSING:
            raise KeyError(f"parameter {key!r} not found in list {self.name!r}")
        return default

    def sublist(self, key: str) -> "ParameterList":
        v = self.get(key)
        if not isinstance(v, ParameterList):
            raise KeyError(f"{key!r} in list {self.name!r} is not a sublist")
        return v

    def isSublist(self, key: str) -> bool:  # noqa: N802 - Teuchos name
        return isinstance(dict.get(self, key), ParameterList)

    def isParameter(self, key: str) -> bool:  # noqa: N802
        return key in self


def _array(text: str, conv):
    t = text.strip()
    if not (t.startswith("{") and t.endswith("}")):
        raise ValueError(f"array value must be written {{a, b, ...}}, got {text!r}")
    body = t[1:-1].strip()
    return tuple(conv(v.strip()) for v in body.split(",")) if body else ()


def _bool(text: str) -> bool:
    t = text.strip().lower()
    if t in ("true", "1", "yes"):
        return True
    if t in ("false", "0", "no"):
        return False
    raise ValueError(f"not a bool: {text!r}")


_TYPES = {
    "double": float, "float": float, "int": int, "unsigned": int, "bool": _bool, "string": str,
    "Array(double)": lambda s: _array(s, float), "Array(int)": lambda s: _array(s, int),
    "Array(string)": lambda s: _array(s, str),
}


def _parse(elem) -> ParameterList:
    out = ParameterList(elem.get("name", ""))
    for child in elem:
        name = child.get("name")
        if name is None:
            raise ValueError(f"<{child.tag}> without a name in list {out.name!r}")
        if child.tag == "ParameterList":
            out[name] = _parse(child)
        elif child.tag == "Parameter":
            typ, val = child.get("type", "string"), child.get("value")
            if typ not in _TYPES:
                raise ValueError(f"parameter {name!r}: unsupported type {typ!r}")
            if val is None:
                raise ValueError(f"parameter {name!r} has no value")
            out[name] = _TYPES[typ](val)
        else:
            raise ValueError(f"unexpected element <{child.tag}> in list {out.name!r}")
    return out


def parse_parameter_list(text: str) -> ParameterList:
    root = ET.fromstring(text)
    if root.tag != "ParameterList":
        raise ValueError("the root element must be a <ParameterList>")
    return _parse(root)


def read_parameter_list(path: str | Path) -> ParameterList:
    return parse_parameter_list(Path(path).read_text())


# ------------------------------------------------------------------ penalty models
_PENALTY_DEFAULTS = {  # (Exponent, Minimum Value)
    "SIMP": (3.0, 1e-9), "RAMP": (3.0, 0.0), "Heaviside": (1.0, 0.0), "NoPenalty": (1.0, 0.0),
}


def penalty_of(params: ParameterList | None, default_type: str = "SIMP") -> tuple[str, float, float, float]:
    """(Type, Exponent, Minimum Value, Regularization Length) of a "Penalty Function" list."""
    params = params if params is not None else ParameterList("Penalty Function")
    typ = params.get("Type", default_type)
    if typ not in _PENALTY_DEFAULTS:
        raise ValueError(f"Penalty Function Type {typ!r} is not supported ({sorted(_PENALTY_DEFAULTS)})")
    p0, m0 = _PENALTY_DEFAULTS[typ]
    return (typ, float(params.get("Exponent", p0)), float(params.get("Minimum Value", m0)),
            float(params.get("Regularization Length", 1.0)))


class _Criterion:
    def __init__(self, name: str, kind: str, model: ElasticModel, own_handle: bool, exponent: float = 0.0,
                 measure: str = "voigt", relaxation: float = 0.5, volume=None):
        self.name, self.kind, self.model, self.own_handle = name, kind, model, own_handle
        self.exponent, self.measure, self.relaxation, self.volume = exponent, measure, relaxation, volume

    @property
    def linear(self) -> bool:
        return self.kind == "Volume"


class ElasticProblem:
    """Plato::Elliptic::Problem<Mechanics<3>> from a deck ("Plato Problem" list) and a mesh."""

    def __init__(self, params: ParameterList, ex: ExodusMesh, device: str = "cuda"):
        pp = params.sublist("Plato Problem") if params.isSublist("Plato Problem") else params
        physics, pde = pp.get("Physics", "Mechanical"), pp.get("PDE Constraint")
        if physics != "Mechanical" or pde != "Elliptic":
            raise ValueError(f"only Physics 'Mechanical' with PDE Constraint 'Elliptic' is supported, "
                             f"got {physics!r} / {pde!r}")
        self.params, self.mesh, self.device = pp, ex, device
        self.self_adjoint = bool(pp.get("Self-Adjoint", False))
        E, nu = self._material(pp, ex)
        self.E, self.nu = E, nu
        self.penalty = penalty_of(pp.sublist("Elliptic").get("Penalty Function", None)
                                  if pp.isSublist("Elliptic") else None)
        self.bcs = self._essential(pp, ex)
        self.tractions = self._natural(pp, ex)
        ls = pp.sublist("Linear Solver") if pp.isSublist("Linear Solver") else ParameterList("Linear Solver")
        precision = str(ls.get("Precision", "FP64")).lower()
        pre = str(ls.get("Preconditioner", "Jacobi")).lower()
        if precision not in ("fp64", "mixed") or pre not in ("jacobi", "multigrid"):
            raise ValueError(f"Linear Solver: Precision {precision!r} / Preconditioner {pre!r} not supported")
        # a box mesh is declared whatever the solver: it also selects the lattice assembly
        self._pre = pre
        self.box = infer_box(ex.mesh)
        if pre == "multigrid" and self.box is None:
            raise ValueError("Linear Solver: Multigrid needs a structured box mesh (tensor lattice, 6 tets per hex)")
        self.model = self._handle(self.penalty, with_box=True)
        self.problem = Problem(self.model, tolerance=float(ls.get("Tolerance", 1e-10)),
                               max_iters=int(ls.get("Iterations", 200000)), precision=precision,
                               preconditioner=pre)
        self.pde = VectorFunction(self.model)
        self.criteria = self._criteria(pp) if pp.isSublist("Criteria") else {}

    # -------------------------------------------------------------- deck sections
    @staticmethod
    def _material(pp: ParameterList, ex: ExodusMesh) -> tuple[float, float]:
        names = set()
        if pp.isSublist("Spatial Model"):
            doms = pp.sublist("Spatial Model").sublist("Domains")
            for dname, d in doms.items():
                blk = d.get("Element Block")
                if blk not in ex.blocks:
                    raise KeyError(f"domain {dname!r}: element block {blk!r} not in the mesh ({sorted(ex.blocks)})")
                names.add(d.get("Material Model"))
        mm = pp.sublist("Material Models")
        if not names:
            names = set(mm.keys())
        if len(names) != 1:
            raise ValueError(f"exactly one material model is supported, the deck names {sorted(names)}")
        name = names.pop()
        m = mm.sublist(name)
        if not m.isSublist("Isotropic Linear Elastic"):
            raise ValueError(f"material {name!r}: only 'Isotropic Linear Elastic' is supported")
        iso = m.sublist("Isotropic Linear Elastic")
        return float(iso.get("Youngs Modulus")), float(iso.get("Poissons Ratio"))

    @staticmethod
    def _essential(pp: ParameterList, ex: ExodusMesh) -> list[EssentialBC]:
        out = []
        if not pp.isSublist("Essential Boundary Conditions"):
            return out
        for name, bc in pp.sublist("Essential Boundary Conditions").items():
            if not isinstance(bc, ParameterList):
                raise ValueError("Essential Boundary Conditions: expect lists only")
            typ = bc.get("Type")
            if typ == "Zero Value":
                value = 0.0
            elif typ == "Fixed Value":
                value = float(bc.get("Value"))
            else:
                raise ValueError(f"essential BC {name!r}: type {typ!r} is not 'Zero Value' or 'Fixed Value'")
            sides = bc.get("Sides")
            if sides not in ex.node_sets:
                raise KeyError(f"essential BC {name!r}: node set {sides!r} not in the mesh ({sorted(ex.node_sets)})")
            dof = int(bc.get("Index", 0))
            if not 0 <= dof <= 2:
                raise ValueError(f"essential BC {name!r}: Index {dof} out of range")
            out.append(EssentialBC(ex.node_sets[sides], dof, value))
        return out

    @staticmethod
    def _natural(pp: ParameterList, ex: ExodusMesh) -> list[Traction]:
        out = []
        if not pp.isSublist("Natural Boundary Conditions"):
            return out
        for name, bc in pp.sublist("Natural Boundary Conditions").items():
            if not isinstance(bc, ParameterList):
                raise ValueError("Natural Boundary Conditions: expect lists only")
            typ = bc.get("Type")
            if typ != "Uniform":
                raise ValueError(f"natural BC {name!r}: type {typ!r} is not supported (Uniform)")
            has_values, has_value = bc.isParameter("Values"), bc.isParameter("Value")
            if has_values == has_value:
                raise ValueError(f"natural BC {name!r}: give exactly one of 'Values' and 'Value'")
            if has_values:
                t = tuple(float(v) for v in bc.get("Values"))
                if len(t) != 3:
                    raise ValueError(f"natural BC {name!r}: 'Values' needs 3 components")
            else:
                t = [0.0, 0.0, 0.0]
                t[int(bc.get("Index", 0))] = float(bc.get("Value"))
                t = tuple(t)
            sides = bc.get("Sides")
            if sides not in ex.side_sets:
                raise KeyError(f"natural BC {name!r}: side set {sides!r} not in the mesh ({sorted(ex.side_sets)})")
            fv, fe = ex.side_sets[sides]
            out.append(Traction(fv, fe, t))
        return out

    def _handle(self, penalty, with_box: bool = False) -> ElasticModel:
        typ, p, rmin, length = penalty

        box = self.box if with_box else None
        # a tensor lattice with unequal spacings is no box: only the multigrid solve needs
        # one, otherwise the same handle goes on with the general assembly
        m = ElasticModel(self.mesh.mesh, self.E, self.nu, p, rmin, self.bcs, self.tractions,
                         device=self.device, box=box, penalty_model=typ, regularization_length=length,
                         box_optional=self._pre != "multigrid")
        if box is not None and m.box is None:
            self.box = None
        return m

    def _criteria(self, pp: ParameterList) -> dict:
        out = {}
        for name, c in pp.sublist("Criteria").items():
            if not isinstance(c, ParameterList):
                raise ValueError("Criteria: expect lists only")
            if bool(c.get("Linear", False)):
                ftype = c.get("Scalar Function Type", "")
                if ftype != "Volume":
                    raise ValueError(f"criterion {name!r}: linear scalar function {ftype!r} is not supported (Volume)")
                typ, p, rmin, length = penalty_of(c.get("Penalty Function", None), default_type="NoPenalty")
                out[name] = _Criterion(name, "Volume", self.model, False,
                                       volume=Volume(self.model, typ, p, rmin, length))
                continue
            if c.get("Type", "Scalar Function") != "Scalar Function":
                raise ValueError(f"criterion {name!r}: Type {c.get('Type')!r} is not supported (Scalar Function)")
            ftype = c.get("Scalar Function Type")
            pen = penalty_of(c.get("Penalty Function", None))
            own = pen != self.penalty
            model = self._handle(pen) if own else self.model
            if ftype == "Internal Elastic Energy":
                out[name] = _Criterion(name, "Energy", model, own)
            elif ftype == "Stress P-Norm":
                measure = str(c.get("Measure", "Voigt"))
                if measure not in ("Voigt", "Relaxed Von Mises"):
                    raise ValueError(f"criterion {name!r}: Measure {measure!r} (Voigt | Relaxed Von Mises)")
                out[name] = _Criterion(name, "PNorm", model, own, exponent=float(c.get("Exponent")),
                                       measure="voigt" if measure == "Voigt" else "von_mises",
                                       relaxation=float(c.get("Relaxation Exponent", 0.5)))
            else:
                raise ValueError(f"criterion {name!r}: scalar function type {ftype!r} is not supported")
        return out

    # -------------------------------------------------------------- Problem interface
    def criterion(self, name: str) -> _Criterion:
        if name not in self.criteria:
            raise KeyError(f"REQUESTED CRITERION NOT DEFINED BY USER: {name!r}")
        return self.criteria[name]

    def criterionIsLinear(self, name: str) -> bool:  # noqa: N802
        return self.criterion(name).linear

    def solution(self, z: torch.Tensor) -> torch.Tensor:
        return self.problem.solution(z)

    def _pnorm(self, c: _Criterion):
        from .criteria import StressPNorm
        return StressPNorm(c.model, c.exponent, c.measure, c.relaxation)

    def criterionValue(self, z: torch.Tensor, u: torch.Tensor, name: str) -> float:  # noqa: N802
        c = self.criterion(name)
        if c.kind == "Volume":
            return c.volume.value(z)
        if c.kind == "Energy":
            return PhysicsScalarFunction(c.model).value(u, z)
        return self._pnorm(c).value(u, z)

    def _partials(self, c: _Criterion, u, z, wrt: str):
        """(dF/du, dF/dz or dF/dX) of a nonlinear criterion at fixed u."""
        if c.kind == "Energy":
            f = PhysicsScalarFunction(c.model)
            return f.gradient_u(u, z), (f.gradient_z(u, z) if wrt == "z" else f.gradient_x(u, z))
        f = self._pnorm(c)
        return f.gradient_u(u, z), (f.gradient_z(u, z) if wrt == "z" else f.gradient_x(u, z))

    def _adjoint(self, u, z, dfdu: torch.Tensor, bc_scale: float) -> torch.Tensor:
        """lambda with K(z) lambda = -dF/du, the residual's K constrained as the reference
        does: criterionGradient with zero values (applyAdjointConstraints),
        criterionGradientX with the prescribed values (applyStateConstraints, scale 1)."""
        K = self.pde.gradient_u_T(u, z).entries
        rhs = dfdu.neg_()
        apply_constraints(self.model, K, rhs, bc_scale)
        lam, st = solve(self.model, K, rhs, tolerance=self.problem.opts.tolerance,
                        max_iters=self.problem.opts.max_iters)
        if not st.converged:
            raise RuntimeError(f"adjoint solve did not converge ({st.iterations} iterations)")
        return lam

    def criterionGradient(self, z: torch.Tensor, u: torch.Tensor, name: str) -> torch.Tensor:  # noqa: N802
        """Problem::criterionGradient (src/elliptic/Problem.hpp:413-468)."""
        from . import _capi
        from .elliptic import _dptr
        c = self.criterion(name)
        if c.kind == "Volume":
            return c.volume.gradient_z(z)
        if self.self_adjoint:
            _, dfdz = self._partials(c, u, z, "z")
            return dfdz.neg_()
        if c.kind == "PNorm" and not c.own_handle:  # fused adjoint with the problem's own solver
            g, _, _ = self._pnorm(c).total_gradient(u, z, self.problem)
            return g
        dfdu, dfdz = self._partials(c, u, z, "z")
        lam = self._adjoint(u, z, dfdu, 0.0)
        dgdz = self.pde.gradient_z(u, z)
        self.model._stream()
        _capi.call("pa_matvec_plus_vector", self.model.handle, 1, 1.0, _dptr(dgdz.entries), _dptr(lam), _dptr(dfdz))
        return dfdz

    def criterionGradientX(self, z: torch.Tensor, u: torch.Tensor, name: str) -> torch.Tensor:  # noqa: N802
        """Problem::criterionGradientX (src/elliptic/Problem.hpp:478-560)."""
        from . import _capi
        from .elliptic import _dptr
        c = self.criterion(name)
        if c.kind == "Volume":
            return c.volume.gradient_x(z)
        if self.self_adjoint:
            _, dfdx = self._partials(c, u, z, "x")
            return dfdx.neg_()
        dfdu, dfdx = self._partials(c, u, z, "x")
        lam = self._adjoint(u, z, dfdu, 1.0)
        dgdx = self.pde.gradient_x(u, z)
        self.model._stream()
        _capi.call("pa_matvec_plus_vector", self.model.handle, 3, 1.0, _dptr(dgdx.entries), _dptr(lam), _dptr(dfdx))
        return dfdx

    def updateProblem(self, z: torch.Tensor, u: torch.Tensor | None) -> None:  # noqa: N802
        """Problem::updateProblem: the supported criteria keep no state between designs
        (their AbstractScalarFunction::updateProblem is the base class's no-op)."""
