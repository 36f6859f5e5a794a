"""pa_solve_noalloc (the block-Jacobi PCG on the library's fixed device workspace) against
pa_solve (the same PCG on handle memory), through the C ABI: two calls in a row, the
second on a smaller mesh, each checked against pa_solve on the same constrained system
and against the true residual; a mesh above the workspace limit is refused before any
device work.  FP64: 1e-12 relative to the largest displacement."""
import ctypes as C

import numpy as np
import pytest
import torch

from platoanalyze_amd import _capi
from platoanalyze_amd import workloads as wl
from platoanalyze_amd.elliptic import ElasticModel, VectorFunction, apply_constraints, solve
from platoanalyze_amd.mesh import build_box

pytestmark = pytest.mark.gpu
MAX_NODES = 131072  # PA_NOALLOC_MAX_NODES in include/plato_mi355x.h


def _system(tiny, seed):
    w = wl.make("cantilever_20x10x10", tiny=tiny)
    m = wl.model(w)
    z = torch.as_tensor(np.random.default_rng(seed).uniform(0.2, 1.0, w.mesh.num_nodes), device="cuda")
    u0 = torch.zeros(m.num_dofs, dtype=torch.float64, device="cuda")
    R = torch.empty_like(u0)
    K = VectorFunction(m).gradient_u(u0, z, residual_out=R)
    b = -R  # R(0, z) = -f
    apply_constraints(m, K.entries, b, 1.0)
    return m, K.entries, b


def _noalloc(m, vals, b, x0=None, tol=1e-10):
    x = x0.clone() if x0 is not None else torch.empty_like(b)
    opts = _capi.SolverOpts(tol, 200000, 16, 0, 1 if x0 is not None else 0, 1e-4, 0)
    st = _capi.SolverStats()
    m._stream()
    _capi.call("pa_solve_noalloc", m.handle, C.c_void_p(vals.data_ptr()), C.c_void_p(x.data_ptr()),
               C.c_void_p(b.data_ptr()), C.byref(opts), C.byref(st))
    return x, st


def _check(m, vals, b, x0=None):
    xr, sr = solve(m, vals, b, tolerance=1e-10, x0=x0)
    xn, sn = _noalloc(m, vals, b, x0)
    torch.cuda.synchronize()
    xr_h, xn_h = xr.cpu().numpy(), xn.cpu().numpy()
    err = np.abs(xn_h - xr_h).max() / np.abs(xr_h).max()
    print(f"  nodes {m.num_nodes}: iterations {sn.iterations} (pa_solve {sr.iterations}), "
          f"rel residual {sn.rel_residual:.2e}, max rel diff {err:.1e}", flush=True)
    assert sn.converged and sn.rel_residual <= 1e-9
    assert sn.iterations == sr.iterations
    assert err <= 1e-12
    return xn


def test_gpu_solve_noalloc_matches_pa_solve():
    print("case: pa_solve_noalloc 20x10x10 then 8x4x4, then a warm start", flush=True)
    m1, v1, b1 = _system(None, 11)
    assert m1.num_nodes <= MAX_NODES
    _check(m1, v1, b1)
    m2, v2, b2 = _system((8, 4, 4), 12)  # second call, smaller input: the shared workspace holds the first's data
    x2 = _check(m2, v2, b2)
    # warm start from a perturbed solution
    x0 = x2 + 1e-3 * torch.randn_like(x2) * x2.abs().max()
    _check(m2, v2, b2, x0=x0)


def test_gpu_solve_noalloc_refuses_large_mesh():
    print("case: pa_solve_noalloc above its node limit", flush=True)
    mesh = build_box(80, 40, 40, 2.0, 1.0, 1.0)  # 136,161 nodes
    assert mesh.num_nodes > MAX_NODES
    m = ElasticModel(mesh, 1.0, 0.3, 3.0, 1e-3)
    n = m.num_dofs
    vals = torch.zeros(9 * m.nnz_blocks, dtype=torch.float64, device="cuda")
    b = torch.ones(n, dtype=torch.float64, device="cuda")
    with pytest.raises(_capi.PaError) as e:
        _noalloc(m, vals, b)
    assert e.value.status == 6 and "exceed" in str(e.value)
