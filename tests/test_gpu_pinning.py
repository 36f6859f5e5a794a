"""GPU checks that close the oracle's pinning through the C ABI: reference goldens the
other suites do not use (Volume3D, the StressPNorm gradient_x), the penalty weights at
interior densities by hand-computed closed forms, the p-norm under every penalty model
against the oracle, an Exodus-built problem against the oracle evaluated on the mesh and
sets read back from the file, and edge cases of the solver and the box declaration.

Tolerances: goldens at the reference's own 1e-13 (1e-12 where its gold is printed
short); oracle parity at 1e-12 relative to the vector's largest magnitude; closed forms
at 1e-13 of the compared quantity.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from platoanalyze_amd import _capi
from platoanalyze_amd import workloads as wl
from platoanalyze_amd.criteria import StressPNorm
from platoanalyze_amd.elliptic import (ElasticModel, EssentialBC, PhysicsScalarFunction, Problem, Traction,
                                       VectorFunction, Volume)
from platoanalyze_amd.mesh import TetMesh, build_box, faces_on_plane

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def close(got, ref, rtol=RTOL, what=""):
    got, ref = np.asarray(got), np.asarray(ref)
    assert got.shape == ref.shape, (what, got.shape, ref.shape)
    err = np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300)
    assert err <= rtol, f"{what}: max error {err:.3e}"


def golden_close(got, gold, rtol):
    """TEST_FLOATING_EQUALITY per entry; golden zeros against the vector's scale."""
    got, gold = np.asarray(got, float), np.asarray(gold, float)
    scale = np.abs(gold).max()
    for g, r in zip(got, gold):
        if r == 0.0:
            assert abs(g) <= 1e-12 * max(scale, 1.0), (g, r)
        else:
            assert abs(g - r) <= rtol * abs(r), (g, r)


def ref_state(nn):
    u = np.empty(3 * nn)
    d = 0.0
    for i in range(3 * nn):
        d += 1e-4
        u[i] = d
    return u


# ------------------------------------------------------------------ reference goldens
def test_gpu_volume_golden_2x2x2(box222, goldens):
    """Volume3D (DerivativeTests.cpp:2151-2302): linear SIMP, z = 1: value 1, dV/dz (27)
    and dV/dX (81) at 1e-13 through pa_volume."""
    print("case: volume golden", flush=True)
    g = goldens["volume3d"]
    m = ElasticModel(box222, 1e6, 0.3, 1.0, 0.0)
    vol = Volume(m, "SIMP", 1.0, 0.0)
    z = dev(np.ones(27))
    assert abs(vol.value(z) - g["value_gold"]) <= 1e-13
    golden_close(host(vol.gradient_z(z)), g["grad_z_gold"], 1e-13)
    golden_close(host(vol.gradient_x(z)), g["grad_x_gold"], 1e-13)


def test_gpu_stress_pnorm_gradient_x_golden_2x2x2(box222, goldens):
    """StressPNorm3D gradient_x (DerivativeTests.cpp:1113-1137), p = 12, 27 values."""
    print("case: p-norm gradient_x golden", flush=True)
    m = ElasticModel(box222, 1e6, 0.3, 1.0, 0.0)
    gx = host(StressPNorm(m, 12.0, "voigt").gradient_x(dev(ref_state(27)), dev(np.ones(27))))
    golden_close(gx[:27], goldens["stress_pnorm3d"]["grad_x_gold"], 1e-12)


@pytest.mark.parametrize("measure,pn,q", [("voigt", 8.0, 0.0), ("von_mises", 8.0, 0.5)])
def test_gpu_stress_pnorm_gradient_x_cfg0(measure, pn, q):
    """dF/dX of both p-norm measures against the oracle's dual-number evaluation on the
    20x10x10 workload with random state and densities."""
    print(f"case: p-norm gradient_x cfg0 {measure}", flush=True)
    w = wl.make("cantilever_20x10x10")
    m = wl.model(w)
    o = orc.Model(w.mesh.coords, w.mesh.conn, w.cfg.E, w.cfg.nu, w.cfg.p, w.cfg.rho_min)
    rng = np.random.default_rng(17)
    u = rng.normal(size=3 * w.mesh.num_nodes) * 1e-2
    z = rng.uniform(1e-3, 1.0, w.mesh.num_nodes)
    F, gx = o.pnorm_gradient_x(u, z, 0 if measure == "voigt" else 1, pn, q)
    close(host(StressPNorm(m, pn, measure, q).gradient_x(dev(u), dev(z))), gx, RTOL, "dF/dX")


# ------------------------------------------------------------------ penalty weights
_S37, _C37 = np.sin(np.radians(37.5)), np.cos(np.radians(37.5))
_H37 = 0.5 * (1.0 + _S37)
# hand-computed w(0.5), w'(0.5): src/Simp.hpp:130-134, src/Ramp.hpp:94-100, src/Heaviside.hpp:110-125
WEIGHTS_AT_HALF = [
    ("SIMP", 3.0, 1e-3, 1.0, 0.125875, 0.74925),
    ("RAMP", 3.0, 1e-3, 1.0, 0.2008, 0.63936),
    ("Heaviside", 2.0, 1e-3, 1.2, 1e-3 + 0.999 * _H37 ** 2, 0.999 * 2.0 * _H37 * 0.5 * _C37 * np.pi / 2.4),
]


@pytest.mark.parametrize("name,p,rmin,L,w,dw", WEIGHTS_AT_HALF)
def test_gpu_penalty_weight_closed_forms(name, p, rmin, L, w, dw):
    """Uniform z = 0.5 on the GPU against the unit-weight problem (penalty "none"):
    K = w K1, R + f = w (R1 + f), C = w C1, sum dC/dz = w' C1, the Voigt p-norm scales by
    w, the volume by w; through both assembly kernels (lattice and general)."""
    print(f"case: closed-form weight {name}", flush=True)
    mesh = build_box(8, 4, 4, 2.0, 1.0, 1.0)
    fv, fe = faces_on_plane(mesh, 0, 2.0)
    fixed = np.flatnonzero(mesh.coords[:, 0] == 0.0).astype(np.int32)
    bcs = [EssentialBC(fixed, d, 0.0) for d in range(3)]
    tr = [Traction(fv, fe, (0.0, 0.0, -1.0))]
    u = dev(np.random.default_rng(3).normal(size=3 * mesh.num_nodes) * 1e-2)
    half, one = dev(np.full(mesh.num_nodes, 0.5)), dev(np.ones(mesh.num_nodes))
    unit = ElasticModel(mesh, 1.0, 0.3, 1.0, 0.0, bcs, tr, penalty_model="NoPenalty")
    K1 = host(VectorFunction(unit).gradient_u(u, one).entries)
    R1 = host(VectorFunction(unit).value(u, one))
    f = -host(VectorFunction(unit).value(torch.zeros_like(u), one))
    C1 = PhysicsScalarFunction(unit).value(u, one)
    F1 = StressPNorm(unit, 8.0, "voigt").value(u, one)
    for box in (None, (8, 4, 4)):
        m = ElasticModel(mesh, 1.0, 0.3, p, rmin, bcs, tr, box=box, penalty_model=name, regularization_length=L)
        R = torch.empty_like(u)
        K = host(VectorFunction(m).gradient_u(u, half, residual_out=R).entries)
        close(K, w * K1, 1e-13, f"K {name} box={box}")
        close(host(R) + f, w * (R1 + f), 1e-13, f"R {name}")
        sf = PhysicsScalarFunction(m)
        assert abs(sf.value(u, half) - w * C1) <= 1e-13 * C1
        assert abs(float(host(sf.gradient_z(u, half)).sum()) - dw * C1) <= 1e-12 * dw * C1
        assert abs(StressPNorm(m, 8.0, "voigt").value(u, half) - w * F1) <= 1e-13 * F1
        assert abs(Volume(m, name, p, rmin, L).value(half) - 2.0 * w) <= 1e-13 * 2.0


@pytest.mark.parametrize("name,kind,p,rmin,L", [("SIMP", 0, 3.0, 1e-3, 1.0), ("RAMP", 1, 3.0, 1e-3, 1.0),
                                                ("Heaviside", 2, 2.0, 1e-3, 1.2), ("NoPenalty", 3, 1.0, 0.0, 1.0)])
def test_gpu_stress_pnorm_penalty_models_cfg0(name, kind, p, rmin, L):
    """The Voigt stress p-norm takes its criterion's own penalty (src/Mechanics.hpp:241-276):
    value, dF/du, dF/dz and dF/dX against the oracle under each model."""
    print(f"case: p-norm penalty {name}", flush=True)
    w = wl.make("cantilever_20x10x10", tiny=(10, 5, 5))
    m = ElasticModel(w.mesh, 1.0, 0.3, p, rmin, penalty_model=name, regularization_length=L)
    o = orc.Model(w.mesh.coords, w.mesh.conn, 1.0, 0.3, p, rmin, penalty_model=kind, length=L)
    rng = np.random.default_rng(23)
    u = rng.normal(size=3 * w.mesh.num_nodes) * 1e-2
    z = rng.uniform(0.05, 1.0, w.mesh.num_nodes)
    du, dz = dev(u), dev(z)
    sp = StressPNorm(m, 8.0, "voigt")
    F, gu, gz = o.pnorm(u, z, 0, 8.0, 0.0)
    assert abs(sp.value(du, dz) - F) <= RTOL * F
    close(host(sp.gradient_u(du, dz)), gu, RTOL, "dF/du")
    if kind != 3:
        close(host(sp.gradient_z(du, dz)), gz, RTOL, "dF/dz")
    close(host(sp.gradient_x(du, dz)), o.pnorm_gradient_x(u, z, 0, 8.0)[1], RTOL, "dF/dX")


# ------------------------------------------------------------------ Exodus ingest vs oracle
def test_gpu_model_from_exodus_vs_oracle(tmp_path):
    """A problem built from an Exodus file's named sets against the oracle evaluated on the
    mesh and sets read back from the file: CRS graph bit-exact, residual (with the
    traction of the side set), Jacobian, the constrained solve against the oracle's PCG
    on the same constrained system, compliance and dC/dz."""
    print("case: exodus vs oracle", flush=True)
    from platoanalyze_amd.exodus import model_from_exodus, read_exodus, sides_of_faces, write_exodus

    w = wl.make("cantilever_20x10x10", tiny=(8, 4, 4))
    el, sd = sides_of_faces(w.mesh, w.face_verts, w.face_elems)
    path = tmp_path / "cantilever.exo"
    write_exodus(path, w.mesh, {"x-": w.fixed_nodes}, {"x+": (el, sd)})
    ex = read_exodus(path)
    m = model_from_exodus(ex, w.cfg.E, w.cfg.nu, w.cfg.p, w.cfg.rho_min,
                          essential={"x-": [(0, 0.0), (1, 0.0), (2, 0.0)]}, tractions={"x+": w.traction})
    fv, fe = ex.side_sets["x+"]
    fixed = ex.node_sets["x-"]
    o = orc.Model(ex.mesh.coords, ex.mesh.conn, w.cfg.E, w.cfg.nu, w.cfg.p, w.cfg.rho_min, fv, w.traction)
    rm, ci = orc.graph(ex.mesh.conn, ex.mesh.num_nodes)
    np.testing.assert_array_equal(host(m.row_map), rm)
    np.testing.assert_array_equal(host(m.column_indices), ci)
    rng = np.random.default_rng(2)
    u = rng.normal(size=3 * ex.mesh.num_nodes) * 1e-2
    z = np.clip(w.z + 0.3 * rng.uniform(-1, 1, ex.mesh.num_nodes), 0.05, 1.0)
    du, dz = dev(u), dev(z)
    R = torch.empty_like(du)
    K = VectorFunction(m).gradient_u(du, dz, residual_out=R)
    close(host(R), o.residual(u, z), RTOL, "residual")
    close(host(K.entries), o.jacobian(u, z, rm, ci), RTOL, "jacobian")
    dofs = np.sort(np.concatenate([3 * fixed + d for d in range(3)])).astype(np.int32)
    f = -o.residual(np.zeros_like(u), z)
    Kc, bc = orc.apply_constraints(rm, ci, o.jacobian(None, z, rm, ci), f, dofs, np.zeros(len(dofs)))
    uo, _, _ = orc.pcg(rm, ci, Kc, bc, tol=1e-13)
    pb = Problem(m, tolerance=1e-12)
    uh = host(pb.solution(dz))
    close(uh, uo, 1e-9, "displacement")
    g, Cv = pb.criterionGradient(dz, dev(uh))
    v, _, gz = o.energy(uh, z, False, True)
    assert abs(Cv - v) <= RTOL * v
    close(host(g), -gz, RTOL, "dC/dz")


# ------------------------------------------------------------------ solver / box edge cases
@pytest.mark.parametrize("precision,preconditioner", [("fp64", "jacobi"), ("mixed", "jacobi"),
                                                      ("fp64", "multigrid")])
def test_gpu_zero_load_warm_start_returns_zero(precision, preconditioner):
    """With no load and zero prescribed values the solution is 0: a nonzero warm start
    (e.g. the previous design's solution) is not returned as converged."""
    print(f"case: zero load warm start {precision} {preconditioner}", flush=True)
    mesh = build_box(8, 4, 4, 2.0, 1.0, 1.0)
    fixed = np.flatnonzero(mesh.coords[:, 0] == 0.0).astype(np.int32)
    m = ElasticModel(mesh, 1.0, 0.3, 3.0, 1e-3, [EssentialBC(fixed, d, 0.0) for d in range(3)], [], box=(8, 4, 4))
    z = dev(np.full(mesh.num_nodes, 0.5))
    pb = Problem(m, tolerance=1e-10, precision=precision, preconditioner=preconditioner)
    u = host(pb.solution(z, u0=dev(np.random.default_rng(1).normal(size=3 * mesh.num_nodes))))
    assert pb.stats.converged and np.abs(u).max() == 0.0


def test_gpu_linspace_axis_keeps_lattice_assembly():
    """A 1000-hex axis laid out by numpy.linspace (nodes off the exact lattice by ~1e-13
    spacings) keeps the lattice assembly, whose values match the general kernel."""
    print("case: linspace axis lattice", flush=True)
    mesh = build_box(1000, 1, 1, 1.0, 0.001, 0.001)
    X = mesh.coords.copy()
    xs = np.linspace(0.0, 1.0, 1001)
    X[:, 0] = xs[np.rint(X[:, 0] * 1000).astype(int)]
    lin = TetMesh(X, mesh.conn)
    mb = ElasticModel(lin, 1.0, 0.3, 3.0, 1e-3, box=(1000, 1, 1))
    v = C.c_int(-1)
    _capi.call("pa_box_assembly", mb.handle, C.byref(v))
    assert v.value == 1
    mg = ElasticModel(lin, 1.0, 0.3, 3.0, 1e-3)
    rng = np.random.default_rng(4)
    u, z = dev(rng.normal(size=3 * lin.num_nodes)), dev(rng.uniform(0.1, 1.0, lin.num_nodes))
    close(host(VectorFunction(mb).gradient_u(u, z).entries), host(VectorFunction(mg).gradient_u(u, z).entries),
          RTOL, "lattice vs general")


@pytest.mark.parametrize("name,p,rmin,L", [("RAMP", 3.0, 1e-3, 1.0), ("Heaviside", 2.0, 1e-3, 1.2),
                                           ("NoPenalty", 1.0, 0.0, 1.0)])
def test_gpu_lattice_fp32_operator_penalty_models(name, p, rmin, L):
    """The mixed solves' lattice FP32 operator takes the penalised element weights of the
    handle's model: under RAMP, Heaviside and no penalty, on a varying density, mixed
    Jacobi and mixed multigrid reach the FP64 Jacobi solution."""
    print(f"case: lattice FP32 operator, penalty {name}", flush=True)
    mesh = build_box(16, 8, 8, 2.0, 1.0, 1.0)
    fv, fe = faces_on_plane(mesh, 0, 2.0)
    fixed = np.flatnonzero(mesh.coords[:, 0] == 0.0).astype(np.int32)
    bcs = [EssentialBC(fixed, d, 0.0) for d in range(3)]
    tr = [Traction(fv, fe, (0.0, 0.0, -1.0))]
    z = dev(np.random.default_rng(8).uniform(0.05, 1.0, mesh.num_nodes))
    m = ElasticModel(mesh, 1.0, 0.3, p, rmin, bcs, tr, box=(16, 8, 8), penalty_model=name,
                     regularization_length=L)
    u64 = host(Problem(m, tolerance=1e-12).solution(z))
    for pre in ("jacobi", "multigrid"):
        pb = Problem(m, tolerance=1e-10, precision="mixed", preconditioner=pre)
        u = host(pb.solution(z))
        err = np.abs(u - u64).max() / np.abs(u64).max()
        print(f"  mixed {pre}: {pb.stats.iterations} it, err {err:.2e}", flush=True)
        assert pb.stats.converged and err <= 1e-7
