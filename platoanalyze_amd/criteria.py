"""Stress p-norm criterion and autograd wrappers over the hand-written adjoint kernels.

* ``StressPNorm``: PhysicsScalarFunction-style value / gradient_u / gradient_z of
  F = (sum_e vol_e m_e^p)^(1/p) -- measure "voigt" is the reference "Stress P-Norm"
  (src/elliptic/StressPNorm.hpp:88-142), measure "von_mises" the relaxed von Mises stress
  rho^q vm(sigma0).
* ``compliance(z, problem)`` and ``stress_pnorm(z, problem, ...)``: torch.autograd
  Functions.  Forward solves K(z) u = f; backward applies the hand-written sensitivity:
  dC/dz = -dC/dz|_u (self-adjoint), dF/dz = dF/dz|_u - (dR/dz)^T lambda with K lambda = dF/du.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _capi
from .elliptic import Problem, _dptr, _dvec

MEASURES = {"voigt": 0, "von_mises": 1}


class StressPNorm:
    def __init__(self, model, exponent: float = 8.0, measure: str = "von_mises", relaxation: float = 0.5):
        if measure not in MEASURES:
            raise ValueError(f"measure must be one of {sorted(MEASURES)}")
        self.m = model
        self.p = float(exponent)
        self.measure = MEASURES[measure]
        self.q = float(relaxation)

    def value(self, u, z) -> float:
        v = C.c_double()
        up, zp = self.m._uz(u, z)
        self.m._stream()
        _capi.call("pa_stress_pnorm", self.m.handle, up, zp, self.measure, self.p, self.q,
                   C.byref(v), None, None)
        return v.value

    def gradient_u(self, u, z):
        g = torch.empty(self.m.num_dofs, dtype=torch.float64, device=u.device)
        up, zp = self.m._uz(u, z)
        self.m._stream()
        _capi.call("pa_stress_pnorm", self.m.handle, up, zp, self.measure, self.p, self.q, None,
                   _dptr(g), None)
        return g

    def gradient_z(self, u, z):
        g = torch.empty(self.m.num_nodes, dtype=torch.float64, device=u.device)
        up, zp = self.m._uz(u, z)
        self.m._stream()
        _capi.call("pa_stress_pnorm", self.m.handle, up, zp, self.measure, self.p, self.q, None,
                   None, _dptr(g))
        return g

    def gradient_x(self, u, z):
        """Partial dF/dX at fixed u and z (3N): PhysicsScalarFunction::gradient_x
        (src/elliptic/PhysicsScalarFunction.hpp:293-355) of the p-norm."""
        g = torch.empty(self.m.num_dofs, dtype=torch.float64, device=u.device)
        up, zp = self.m._uz(u, z)
        self.m._stream()
        _capi.call("pa_stress_pnorm_gradient_x", self.m.handle, up, zp, self.measure, self.p, self.q,
                   _dptr(g), None)
        return g

    def total_gradient(self, u, z, problem: Problem):
        """dF/dz through the state equation; returns (dF/dz, F, lambda)."""
        g = torch.empty(self.m.num_nodes, dtype=torch.float64, device=u.device)
        lam = torch.empty(self.m.num_dofs, dtype=torch.float64, device=u.device)
        v = C.c_double()
        up, zp = self.m._uz(u, z)
        self.m._stream()
        _capi.call("pa_stress_pnorm_gradient", self.m.handle, up, zp, self.measure, self.p, self.q,
                   _dptr(g), _dptr(lam), C.byref(v), C.byref(problem.opts), C.byref(problem.stats))
        return g, v.value, lam


def dRdz_T_apply(model, u, lam, z):
    up, zp = model._uz(u, z)
    lp = _dvec(lam, model.num_dofs, "lambda")
    out = torch.empty(model.num_nodes, dtype=torch.float64, device=u.device)
    model._stream()
    _capi.call("pa_dRdz_T_apply", model.handle, up, lp, zp, _dptr(out))
    return out


class _Compliance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, z, problem):
        z = z.detach().contiguous()
        u = problem.solution(z)
        g, C_val = problem.criterionGradient(z, u)
        ctx.save_for_backward(g)
        return torch.tensor(C_val, dtype=torch.float64, device=z.device)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None


class _StressPNormFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, z, problem, criterion):
        z = z.detach().contiguous()
        u = problem.solution(z)
        g, F, _ = criterion.total_gradient(u, z, problem)
        ctx.save_for_backward(g)
        return torch.tensor(F, dtype=torch.float64, device=z.device)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None, None


def compliance(z: torch.Tensor, problem: Problem) -> torch.Tensor:
    """C(z) = 1/2 u^T K(z) u at equilibrium, differentiable in z (hand-written adjoint)."""
    return _Compliance.apply(z, problem)


def stress_pnorm(z: torch.Tensor, problem: Problem, exponent: float = 8.0, measure: str = "von_mises",
                 relaxation: float = 0.5) -> torch.Tensor:
    """F(z) at equilibrium, differentiable in z (adjoint solve + hand-written kernels)."""
    return _StressPNormFn.apply(z, problem, StressPNorm(problem.m, exponent, measure, relaxation))
