"""FP32 matrix-free operator on the lattice (k_boxmf_apply) inside the mixed-precision
solves, through the C ABI.

The operator changes the iteration, not the system: each mixed solve is checked
against the FP64 block-CRS Jacobi-PCG solution of the same system (pinned to the
oracle in test_gpu_parity.py) at 1e-7 of max |u| for a 1e-10 relative-residual stop on
the true FP64 residual.  The lattice operator sums in a fixed order, so a repeated solve
gives bit-identical results and the same iteration count; the chunk-based FP32 apply
(PA_BOX_MF=0) solves the same system.  On hexes split into the six main-diagonal
Kuhn tets (every box mesh here) the operator is k_boxmf_apply_kuhn, the tet-by-tet
form; PA_BOX_MF_KUHN=0 selects the block-table kernel k_boxmf_apply, and both are
compared with the FP64 operator after a fixed number of PCG iterations.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from platoanalyze_amd import _capi
from platoanalyze_amd import workloads as wl
from platoanalyze_amd.elliptic import Problem, _dptr

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


def _pcg_iterate(m, dz, precision, iters):
    """u after exactly `iters` Jacobi-PCG iterations from u = 0 (the solve stops at its
    iteration cap, PA_ERR_NOT_CONVERGED; mixed precision with no residual replacement in
    between).  It depends on the operator through K p of `iters` directions."""
    u = torch.zeros(m.num_dofs, dtype=torch.float64, device="cuda")
    vals = torch.empty(9 * m.nnz_blocks, dtype=torch.float64, device="cuda") if precision == 0 else None
    opts = _capi.SolverOpts(1e-15, iters, iters, precision, 0, 1e-6, 0)
    stats = _capi.SolverStats()
    m._stream()
    with pytest.raises(_capi.PaError) as e:
        _capi.call("pa_solution", m.handle, _dptr(dz), _dptr(u), None if vals is None else _dptr(vals),
                   C.byref(opts), C.byref(stats))
    assert e.value.status == 5 and stats.iterations == iters
    return host(u)


@pytest.mark.parametrize("dims", [(20, 10, 10), (33, 17, 9)])
def test_gpu_lattice_kuhn_operator_iterates(dims, monkeypatch):
    """Four PCG iterations with the tet-by-tet FP32 operator, the block-table FP32
    operator and the FP64 CRS operator give the same iterate to FP32 rounding (a wrong
    sign, tet, weight or spacing in either kernel changes it at O(1)); uneven hex counts
    and spacings per axis."""
    print(f"case: Kuhn lattice operator iterates {dims}", flush=True)
    w = wl.make("cantilever_128x64x64", tiny=dims)
    dz = dev(w.z)
    m = wl.model(w)
    assert m.box_assembly
    u64 = _pcg_iterate(m, dz, 0, 4)
    uk = _pcg_iterate(m, dz, 1, 4)
    monkeypatch.setenv("PA_BOX_MF_KUHN", "0")
    mb = wl.model(w)
    ub = _pcg_iterate(mb, dz, 1, 4)
    ek, eb = rel_err(uk, u64), rel_err(ub, u64)
    print(f"  after 4 iterations: Kuhn vs FP64 {ek:.2e}, block table vs FP64 {eb:.2e}", flush=True)
    assert ek <= 1e-5 and eb <= 1e-5


@pytest.mark.parametrize("preconditioner", ["jacobi", "multigrid"])
def test_gpu_lattice_kuhn_operator_solves(preconditioner, monkeypatch):
    """The mixed solves with either lattice operator reach the FP64 solution; the
    tet-by-tet kernel takes at most 10 % more iterations than the block-table one."""
    print(f"case: Kuhn lattice operator solves, {preconditioner}", flush=True)
    w = wl.make("cantilever_128x64x64", tiny=(64, 32, 32))
    dz = dev(w.z)
    m = wl.model(w)
    u64 = host(Problem(m, tolerance=1e-12, preconditioner=preconditioner).solution(dz))
    pk = Problem(m, tolerance=1e-10, precision="mixed", preconditioner=preconditioner)
    uk = host(pk.solution(dz))
    monkeypatch.setenv("PA_BOX_MF_KUHN", "0")
    mb = wl.model(w)
    pb = Problem(mb, tolerance=1e-10, precision="mixed", preconditioner=preconditioner)
    ub = host(pb.solution(dz))
    print(f"  Kuhn {pk.stats.iterations} it err {rel_err(uk, u64):.2e}; block table {pb.stats.iterations} it "
          f"err {rel_err(ub, u64):.2e}", flush=True)
    assert pk.stats.converged and pb.stats.converged
    assert rel_err(uk, u64) <= 1e-7 and rel_err(ub, u64) <= 1e-7
    assert pk.stats.iterations <= 1.1 * pb.stats.iterations
