"""FP32 matrix-free operator on the lattice (k_boxmf_apply) inside the mixed-precision
solves, through the C ABI.

The operator changes the iteration, not the system: each mixed solve is checked
against the FP64 block-CRS Jacobi-PCG solution of the same system (pinned to the
oracle in test_gpu_parity.py) at 1e-7 of max |u| for a 1e-10 relative-residual stop on
the true FP64 residual.  The lattice operator sums in a fixed order, so a repeated solve
gives bit-identical results and the same iteration count; the chunk-based FP32 apply
(PA_BOX_MF=0) solves the same system.
"""
import numpy as np
import pytest
import torch

from platoanalyze_amd import workloads as wl
from platoanalyze_amd.elliptic import Problem

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def rel_err(got, ref):
    return np.abs(got - ref).max() / np.abs(ref).max()


@pytest.mark.parametrize("dims", [(20, 10, 10), (64, 32, 32)])
@pytest.mark.parametrize("preconditioner", ["jacobi", "multigrid"])
def test_gpu_lattice_fp32_operator_solves(dims, preconditioner, monkeypatch):
    print(f"case: lattice FP32 operator, {preconditioner} {dims}", flush=True)
    w = wl.make("cantilever_128x64x64", tiny=dims)
    m = wl.model(w)
    assert m.box_assembly
    dz = dev(w.z)
    ref = Problem(m, tolerance=1e-12, preconditioner=preconditioner)
    u64 = host(ref.solution(dz))
    runs = []
    for _ in range(2):
        pb = Problem(m, tolerance=1e-10, precision="mixed", preconditioner=preconditioner)
        u = host(pb.solution(dz))
        assert pb.stats.converged and pb.stats.rel_residual <= 1e-10
        runs.append((u, pb.stats.iterations, pb.stats.outer_iterations))
    (u, it, outer), (u2, it2, _) = runs
    err = rel_err(u, u64)
    # the chunk-based FP32 apply on a handle without the lattice operator
    monkeypatch.setenv("PA_BOX_MF", "0")
    mc = wl.model(w)
    pc = Problem(mc, tolerance=1e-10, precision="mixed", preconditioner=preconditioner)
    uc = host(pc.solution(dz))
    print(f"  FP64 {ref.stats.iterations} it; lattice FP32 operator {it} it (outer {outer}), err {err:.2e}; "
          f"chunk FP32 apply {pc.stats.iterations} it, err {rel_err(uc, u64):.2e}", flush=True)
    assert err <= 1e-7
    assert it == it2 and np.array_equal(u, u2)  # deterministic
    assert pc.stats.converged and rel_err(uc, u64) <= 1e-7
