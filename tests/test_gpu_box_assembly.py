"""Lattice Jacobian assembly (pa_box.hip) against the general chunk kernel and the oracle.

The lattice kernel applies to declared boxes (pa_set_box) whose hexes all carry the same
six tets; it must give the general kernel's K and R = K u - f to rounding (1e-13 of the
largest entry; the two sum the same element terms in a different order) and the oracle's
at the parity bar (1e-12), for every penalty model, with node numberings other than the
Hilbert one, and it must stand aside where its conditions fail.
"""
import os

import ctypes as C

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from platoanalyze_amd import _capi
from platoanalyze_amd.elliptic import ElasticModel, EssentialBC, Traction, VectorFunction
from platoanalyze_amd.mesh import TetMesh, build_box, faces_on_plane

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def active(m) -> int:
    v = C.c_int(-1)
    _capi.call("pa_box_assembly", m.handle, C.byref(v))
    return v.value


def model(mesh, dims, general=False, **kw):
    fixed = np.flatnonzero(mesh.coords[:, 0] == 0.0).astype(np.int32)
    fv, fe = faces_on_plane(mesh, 0, float(mesh.coords[:, 0].max()))
    bcs = [EssentialBC(fixed, d) for d in range(3)]
    tr = [Traction(fv, fe, (0.1, 0.0, -1.0))]
    old = os.environ.get("PA_BOX_ASSEMBLY")
    if general:
        os.environ["PA_BOX_ASSEMBLY"] = "0"
    try:
        return ElasticModel(mesh, 1.0, 0.3, kw.pop("p", 3.0), kw.pop("rmin", 1e-3), bcs, tr, box=dims, **kw)
    finally:
        if general:
            if old is None:
                del os.environ["PA_BOX_ASSEMBLY"]
            else:
                os.environ["PA_BOX_ASSEMBLY"] = old


def both(m, u, z):
    R = torch.empty(m.num_dofs, dtype=torch.float64, device="cuda")
    K = VectorFunction(m).gradient_u(u, z, residual_out=R)
    K0 = VectorFunction(m).gradient_u(None, z)  # Jacobian alone
    return host(K.entries), host(R), host(K0.entries)


def close(got, ref, tol, what):
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err <= tol, f"{what}: {err:.3e}"


@pytest.mark.parametrize("dims", [(4, 2, 2), (20, 10, 10), (7, 5, 3)])
@pytest.mark.parametrize("penalty", [("SIMP", 3.0, 1e-3), ("SIMP", 2.5, 1e-3), ("RAMP", 2.0, 1e-2), ("Heaviside", 4.0, 0.0)])
@pytest.mark.parametrize("ahead", ["0", "1"])
def test_gpu_box_matches_general_and_oracle(dims, penalty, ahead, monkeypatch):
    """ahead = 1 forces the kernel that large lattices take (gathers one tile ahead, the
    load subtracted after the kernel) onto these small ones."""
    monkeypatch.setenv("PA_BOX_AHEAD", ahead)
    name, p, rmin = penalty
    print(f"case: box assembly {dims} {name} ahead={ahead}", flush=True)
    mesh = build_box(*dims, 2.0, 1.0, 1.0)
    rng = np.random.default_rng(sum(dims))
    z = rng.uniform(0.05, 1.0, mesh.num_nodes)
    u = rng.normal(size=3 * mesh.num_nodes)
    mf = model(mesh, dims, penalty_model=name, p=p, rmin=rmin, regularization_length=0.5)
    mg = model(mesh, dims, general=True, penalty_model=name, p=p, rmin=rmin, regularization_length=0.5)
    assert active(mf) == 1 and active(mg) == 0
    Kf, Rf, K0f = both(mf, dev(u), dev(z))
    Kg, Rg, _ = both(mg, dev(u), dev(z))
    close(Kf, Kg, 1e-13, "K box vs general")
    close(Rf, Rg, 1e-13, "R box vs general")
    np.testing.assert_array_equal(K0f, Kf)  # the Jacobian alone is the same launch's values
    from platoanalyze_amd.elliptic import penalty_model_id
    o = orc.Model(mesh.coords, mesh.conn, 1.0, 0.3, p, rmin, mf.tractions[0].face_verts, (0.1, 0.0, -1.0),
                  penalty_model=penalty_model_id(name), length=0.5)
    rm, ci = host(mf.row_map), host(mf.column_indices)
    close(Kf, o.jacobian(u, z, rm, ci), 1e-12, "K box vs oracle")
    close(Rf, o.residual(u, z), 1e-12, "R box vs oracle")


def test_gpu_box_permuted_nodes():
    """Nodes in a random order (not the Hilbert one) and elements shuffled: the lattice
    kernel still applies and matches the general kernel."""
    print("case: box assembly permuted", flush=True)
    dims = (6, 4, 3)
    base = build_box(*dims, 2.0, 1.0, 1.0)
    rng = np.random.default_rng(3)
    perm = rng.permutation(base.num_nodes)          # new id of old node
    coords = np.empty_like(base.coords)
    coords[perm] = base.coords
    conn = perm[base.conn][rng.permutation(base.num_elems)].astype(np.int32)
    mesh = TetMesh(coords, conn)
    z = rng.uniform(0.1, 1.0, mesh.num_nodes)
    u = rng.normal(size=3 * mesh.num_nodes)
    mf, mg = model(mesh, dims), model(mesh, dims, general=True)
    assert active(mf) == 1
    Kf, Rf, _ = both(mf, dev(u), dev(z))
    Kg, Rg, _ = both(mg, dev(u), dev(z))
    close(Kf, Kg, 1e-13, "K")
    close(Rf, Rg, 1e-13, "R")


def test_gpu_box_stands_aside():
    """Nodes off the exact lattice (by 1e-9 of a spacing, inside the multigrid's lattice
    tolerance) or a box without the same six tets in every hex keep the general kernel."""
    print("case: box assembly conditions", flush=True)
    dims = (4, 2, 2)
    mesh = build_box(*dims, 2.0, 1.0, 1.0)
    moved = TetMesh(mesh.coords.copy(), mesh.conn.copy())
    moved.coords[7, 1] += 1e-9 * 0.5
    assert active(model(moved, dims)) == 0
    # swap the split of one hex: rotate its six tets' vertices by the mirror x -> 1 - x
    m2 = TetMesh(mesh.coords.copy(), mesh.conn.copy())
    X = m2.coords
    h = np.array([0.5, 0.5, 0.5])
    cells = [e for e in range(m2.num_elems) if np.all(X[m2.conn[e]].min(axis=0) == 0.0)]
    assert len(cells) == 6
    lat = {tuple(np.round(X[n] / h).astype(int)): n for n in range(m2.num_nodes)}
    for e in cells:
        for a in range(4):
            i, j, k = np.round(X[m2.conn[e, a]] / h).astype(int)
            m2.conn[e, a] = lat[(1 - i, j, k)]
    assert active(model(m2, dims)) == 0
    # the general kernel still assembles it
    m = model(m2, dims)
    z = dev(np.full(m2.num_nodes, 0.7))
    assert np.isfinite(host(VectorFunction(m).gradient_u(None, z).entries)).all()


@pytest.mark.parametrize("dims", [(4, 2, 2), (20, 10, 10), (7, 5, 3)])
@pytest.mark.parametrize("penalty", [("SIMP", 3.0, 1e-3), ("RAMP", 2.0, 1e-2)])
@pytest.mark.parametrize("kuhn", ["1", "0"])
def test_gpu_box_value_matches_general_and_oracle(dims, penalty, kuhn, monkeypatch):
    """VectorFunction::value on a declared box: the tet-by-tet FP64 residual (default on
    the main-diagonal Kuhn split) or the stencil residual (PA_BOX_RES_KUHN=0) against the
    general kernel and the oracle, uneven spacings per axis."""
    monkeypatch.setenv("PA_BOX_RES_KUHN", kuhn)
    name, p, rmin = penalty
    print(f"case: box value {dims} {name} kuhn={kuhn}", flush=True)
    mesh = build_box(*dims, 2.0, 1.0, 0.7)
    rng = np.random.default_rng(7 + sum(dims))
    z = rng.uniform(0.05, 1.0, mesh.num_nodes)
    u = rng.normal(size=3 * mesh.num_nodes)
    mf = model(mesh, dims, penalty_model=name, p=p, rmin=rmin)
    mg = model(mesh, dims, general=True, penalty_model=name, p=p, rmin=rmin)
    assert active(mf) == 1 and active(mg) == 0
    Rf = host(VectorFunction(mf).value(dev(u), dev(z)))
    Rg = host(VectorFunction(mg).value(dev(u), dev(z)))
    close(Rf, Rg, 1e-13, "R box vs general")
    from platoanalyze_amd.elliptic import penalty_model_id
    o = orc.Model(mesh.coords, mesh.conn, 1.0, 0.3, p, rmin, mf.tractions[0].face_verts, (0.1, 0.0, -1.0),
                  penalty_model=penalty_model_id(name))
    close(Rf, o.residual(u, z), 1e-12, "R box vs oracle")
