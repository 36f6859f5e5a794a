"""HIP path (through the C ABI) against the oracle and the reference goldens.

Tolerance: FP64, 1e-12 relative to the largest magnitude of the compared vector
(entries that are exact zeros in one evaluation order are ~1e-17 in another), except
where the reference's own tighter tolerance is restated (goldens).  The CRS graph is
compared bit-exactly.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from platoanalyze_amd import workloads as wl
from platoanalyze_amd.elliptic import (ElasticModel, EssentialBC, PhysicsScalarFunction, Problem, Traction,
                                       VectorFunction, apply_constraints, solve)
from platoanalyze_amd.mesh import TetMesh, build_box

pytestmark = pytest.mark.gpu
RTOL = 1e-12
REF_ROW0 = [0, 8, 1, 3, 9, 13, 21, 25]


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def close(got, ref, rtol=RTOL, what=""):
    got, ref = np.asarray(got), np.asarray(ref)
    assert got.shape == ref.shape, (what, got.shape, ref.shape)
    scale = max(np.abs(ref).max(), 1e-300)
    err = np.abs(got - ref).max() / scale
    assert err <= rtol, f"{what}: max error {err:.3e} relative to max |ref| {scale:.3e}"


def ref_state(nn):
    u = np.empty(3 * nn)
    d = 0.0
    for i in range(3 * nn):
        d += 1e-4
        u[i] = d
    return u


def oracle_of(w):
    return orc.Model(w.mesh.coords, w.mesh.conn, w.cfg.E, w.cfg.nu, w.cfg.p, w.cfg.rho_min, w.face_verts,
                     w.traction)


# ------------------------------------------------------------------ goldens
def test_gpu_reference_goldens_2x2x2(box222, goldens):
    print("case: 2x2x2 goldens", flush=True)
    rm, ci = orc.graph(box222.conn, 27)
    ci = ci.copy()
    ci[rm[0]:rm[1]] = REF_ROW0  # the reference's Omega_h column order of row 0
    m = ElasticModel(box222, 1e6, 0.3, 1.0, 0.0, graph=(rm, ci))
    vf = VectorFunction(m)
    u, z = dev(ref_state(27)), dev(np.ones(27))
    g = goldens["elastostatic_residual3d"]
    close(host(vf.value(u, z)), g["residual_gold"], 1e-13, "residual")
    K = host(vf.gradient_u(u, z).entries)
    np.testing.assert_allclose(K[:45], g["gold_jac_entries"], rtol=1e-14, atol=1e-9)
    G = host(vf.gradient_z(u, z).entries)
    close(G[:24], g["gold_grad_entries"], 2e-14, "gradient_z")
    np.testing.assert_array_equal(host(m.row_map), rm)
    np.testing.assert_array_equal(host(m.column_indices), ci)
    e = goldens["internal_elastic_energy3d"]
    sf = PhysicsScalarFunction(m)
    assert abs(sf.value(u, z) - e["value_gold"]) <= 1e-13 * e["value_gold"]
    close(host(sf.gradient_u(u, z)), e["grad_u_gold"], 1e-12, "energy grad_u")
    close(host(sf.gradient_z(u, z)), e["grad_z_gold"], 1e-13, "energy grad_z")


# ------------------------------------------------------------------ config 0
@pytest.fixture(scope="module")
def cfg0():
    w = wl.make("cantilever_20x10x10")
    m = wl.model(w)
    o = oracle_of(w)
    rm, ci = orc.graph(w.mesh.conn, w.mesh.num_nodes)
    rng = np.random.default_rng(11)
    u = rng.normal(size=3 * w.mesh.num_nodes) * 1e-2
    z = rng.uniform(1e-3, 1.0, w.mesh.num_nodes)
    return w, m, o, rm, ci, u, z


def test_gpu_graph_bitexact_cfg0(cfg0):
    w, m, o, rm, ci, u, z = cfg0
    np.testing.assert_array_equal(host(m.row_map), rm)
    np.testing.assert_array_equal(host(m.column_indices), ci)


def test_gpu_residual_jacobian_cfg0(cfg0):
    print("case: cfg0 residual/jacobian", flush=True)
    w, m, o, rm, ci, u, z = cfg0
    vf = VectorFunction(m)
    du, dz = dev(u), dev(z)
    close(host(vf.value(du, dz)), o.residual(u, z), RTOL, "residual")
    R2 = torch.empty_like(du)
    K = vf.gradient_u(du, dz, residual_out=R2)
    Ko = o.jacobian(u, z, rm, ci)
    close(host(K.entries), Ko, RTOL, "jacobian")
    close(host(R2), o.residual(u, z), RTOL, "fused residual")
    close(host(vf.gradient_u(None, dz).entries), Ko, RTOL, "jacobian without state")
    close(host(vf.gradient_z(du, dz).entries), o.gradient_z(u, z, rm, ci), RTOL, "gradient_z")


def test_gpu_energy_cfg0(cfg0):
    w, m, o, rm, ci, u, z = cfg0
    sf = PhysicsScalarFunction(m)
    du, dz = dev(u), dev(z)
    v, gu, gz = o.energy(u, z)
    assert abs(sf.value(du, dz) - v) <= RTOL * abs(v)
    close(host(sf.gradient_u(du, dz)), gu, RTOL, "energy grad_u")
    close(host(sf.gradient_z(du, dz)), gz, RTOL, "energy grad_z")


def test_gpu_constraints_cfg0(cfg0):
    w, m, o, rm, ci, u, z = cfg0
    Ko = o.jacobian(None, z, rm, ci)
    b = np.random.default_rng(5).normal(size=3 * w.mesh.num_nodes)
    Kc_o, bc_o = orc.apply_constraints(rm, ci, Ko, b, w.bc_dofs, np.zeros(len(w.bc_dofs)))
    K = VectorFunction(m).gradient_u(None, dev(z))
    db = dev(b)
    apply_constraints(m, K.entries, db, 1.0)
    close(host(K.entries), Kc_o, RTOL, "constrained matrix")
    close(host(db), bc_o, RTOL, "constrained rhs")


def test_gpu_nonzero_dirichlet_values():
    """Prescribed (nonzero) essential values: lifting of the rhs matches the oracle."""
    mesh = build_box(6, 3, 3, 2.0, 1.0, 1.0)
    fixed = np.flatnonzero(mesh.coords[:, 0] == 0.0).astype(np.int32)
    bcs = [EssentialBC(fixed, 0, 0.0), EssentialBC(fixed, 1, 0.01), EssentialBC(fixed, 2, -0.02)]
    m = ElasticModel(mesh, 1.0, 0.3, 3.0, 1e-3, bcs)
    z = np.random.default_rng(2).uniform(0.1, 1.0, mesh.num_nodes)
    rm, ci = orc.graph(mesh.conn, mesh.num_nodes)
    o = orc.Model(mesh.coords, mesh.conn, 1.0, 0.3, 3.0, 1e-3)
    dofs = np.concatenate([3 * fixed, 3 * fixed + 1, 3 * fixed + 2])
    vals = np.concatenate([np.zeros(len(fixed)), np.full(len(fixed), 0.01), np.full(len(fixed), -0.02)])
    b = np.random.default_rng(3).normal(size=3 * mesh.num_nodes)
    Kc_o, bc_o = orc.apply_constraints(rm, ci, o.jacobian(None, z, rm, ci), b, dofs, vals, 0.5)
    K = VectorFunction(m).gradient_u(None, dev(z))
    db = dev(b)
    apply_constraints(m, K.entries, db, 0.5)
    close(host(K.entries), Kc_o, RTOL, "matrix")
    close(host(db), bc_o, RTOL, "rhs")


def test_gpu_solution_and_criterion_cfg0(cfg0):
    print("case: cfg0 solve", flush=True)
    w, m, o, rm, ci, u, z = cfg0
    z = w.z  # uniform 0.5 as the workload specifies
    pb = Problem(m, tolerance=1e-12)
    dz = dev(z)
    du = pb.solution(dz)
    assert pb.stats.converged and pb.stats.rel_residual <= 1e-11
    # oracle: same constrained system, PCG and dense solve agree
    f = -o.residual(np.zeros(3 * w.mesh.num_nodes), z)
    Kc, bc = orc.apply_constraints(rm, ci, o.jacobian(None, z, rm, ci), f, w.bc_dofs, np.zeros(len(w.bc_dofs)))
    uo, it, rel = orc.pcg(rm, ci, Kc, bc, tol=1e-12)
    uh = host(du)
    close(uh, uo, 1e-9, "displacement")  # both solves stop at 1e-12 relative residual
    # true residual of the GPU solution in the oracle's arithmetic
    r = orc.matvec(rm, ci, Kc, uh) - bc
    assert np.linalg.norm(r) <= 1e-11 * np.linalg.norm(bc)
    g, C = pb.criterionGradient(dz, du)
    v, _, gz = o.energy(uh, z, False, True)
    assert abs(C - v) <= RTOL * abs(v)
    assert abs(C - 0.5 * f @ uh) <= 1e-10 * abs(C)  # compliance = 1/2 f.u at equilibrium
    close(host(g), -gz, RTOL, "dC/dz")


def test_gpu_compliance_gradient_fd():
    """Self-adjoint total derivative dC/dz against central differences through full solves."""
    w = wl.make("cantilever_20x10x10", tiny=(8, 4, 4))
    m = wl.model(w)
    z = np.random.default_rng(4).uniform(0.3, 1.0, w.mesh.num_nodes)
    pb = Problem(m, tolerance=1e-13)
    g, C0 = pb.criterionGradient(dev(z))
    g = host(g)
    for n in [3, 40, 77, w.mesh.num_nodes - 5]:
        h = 1e-6
        zp, zm = z.copy(), z.copy()
        zp[n] += h
        zm[n] -= h
        fd = (pb.criterionValue(dev(zp)) - pb.criterionValue(dev(zm))) / (2 * h)
        assert abs(fd - g[n]) <= 1e-6 * np.abs(g).max(), (n, fd, g[n])


# ------------------------------------------------------------------ edge cases
def _random_delaunay(npts, seed):
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    pts = rng.uniform(0, 1, (npts, 3))
    tri = Delaunay(pts)
    conn = tri.simplices.astype(np.int32)
    X = pts[conn]
    J = np.stack([X[:, 0] - X[:, 3], X[:, 1] - X[:, 3], X[:, 2] - X[:, 3]], axis=2)
    keep = np.abs(np.linalg.det(J)) > 1e-6  # drop slivers
    return TetMesh(pts, np.ascontiguousarray(conn[keep]))


@pytest.mark.parametrize("seed", [0, 1])
def test_gpu_unstructured_mesh_wide_rows(seed):
    """Unstructured Delaunay tets (both orientations, rows wider than 16 blocks -> 8-bit
    slot codes, chunks with ragged row counts)."""
    mesh = _random_delaunay(400, seed)
    rm, ci = orc.graph(mesh.conn, mesh.num_nodes)
    assert np.diff(rm).max() > 16
    m = ElasticModel(mesh, 2.0, 0.25, 3.0, 1e-3)
    np.testing.assert_array_equal(host(m.column_indices), ci)
    rng = np.random.default_rng(seed + 10)
    u = rng.normal(size=3 * mesh.num_nodes)
    z = rng.uniform(0.05, 1.0, mesh.num_nodes)
    o = orc.Model(mesh.coords, mesh.conn, 2.0, 0.25, 3.0, 1e-3)
    vf, sf = VectorFunction(m), PhysicsScalarFunction(m)
    du, dz = dev(u), dev(z)
    close(host(vf.value(du, dz)), o.residual(u, z), RTOL, "residual")
    close(host(vf.gradient_u(du, dz).entries), o.jacobian(u, z, rm, ci), RTOL, "jacobian")
    close(host(vf.gradient_z(du, dz).entries), o.gradient_z(u, z, rm, ci), RTOL, "gradient_z")
    v, gu, gz = o.energy(u, z)
    assert abs(sf.value(du, dz) - v) <= RTOL * abs(v)
    close(host(sf.gradient_z(du, dz)), gz, RTOL, "energy grad_z")


def test_gpu_single_element_and_permuted_graph():
    mesh = TetMesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]]), np.array([[0, 2, 1, 3]], np.int32))
    # caller graph with descending columns
    rm = np.array([0, 4, 8, 12, 16], np.int32)
    ci = np.array([3, 2, 1, 0] * 4, np.int32)
    m = ElasticModel(mesh, 1.0, 0.3, 3.0, 1e-9, graph=(rm, ci))
    u = np.arange(12, dtype=float) * 1e-3
    z = np.array([0.2, 0.4, 0.6, 0.8])
    o = orc.Model(mesh.coords, mesh.conn, 1.0, 0.3, 3.0, 1e-9)
    close(host(VectorFunction(m).gradient_u(dev(u), dev(z)).entries), o.jacobian(u, z, rm, ci), RTOL, "K")
    close(host(VectorFunction(m).value(dev(u), dev(z))), o.residual(u, z), RTOL, "R")


def test_gpu_errors_are_raised():
    from platoanalyze_amd._capi import PaError

    w = wl.make("cantilever_20x10x10", tiny=(4, 2, 2))
    m = wl.model(w)
    # singular system (no essential BCs) -> PCG cannot converge within the cap
    free = ElasticModel(w.mesh, 1.0, 0.3, 3.0, 1e-3, [], [Traction(w.face_verts, w.face_elems, (1.0, 0.0, 0.0))])
    pb = Problem(free, tolerance=1e-14, max_iters=50)
    with pytest.raises(PaError):
        pb.solution(dev(w.z))
    with pytest.raises(ValueError):
        VectorFunction(m).value(dev(np.zeros(5)), dev(w.z))
    _ = solve  # exported


# ------------------------------------------------------------------ scale
@pytest.mark.slow
@pytest.mark.parametrize("assembly", ["lattice", "general"])
def test_gpu_scale_cfg1_parity(assembly, monkeypatch):
    """128x64x64 x 6 Tet4 (3.15M tets): graph bit-exact, residual / Jacobian / energy
    gradient against the oracle on the same smoothed densities, with the lattice assembly
    the box declaration enables and with the general chunk kernel (PA_BOX_ASSEMBLY=0, the
    path of every unstructured or file mesh)."""
    print(f"case: cfg1 scale parity, {assembly} assembly", flush=True)
    w = wl.make("cantilever_128x64x64")
    if assembly == "general":
        monkeypatch.setenv("PA_BOX_ASSEMBLY", "0")
    m = wl.model(w)
    assert m.box_assembly == (assembly == "lattice")
    rm, ci = orc.graph(w.mesh.conn, w.mesh.num_nodes)
    np.testing.assert_array_equal(host(m.row_map), rm)
    np.testing.assert_array_equal(host(m.column_indices), ci)
    o = oracle_of(w)
    rng = np.random.default_rng(21)
    u = rng.normal(size=3 * w.mesh.num_nodes) * 1e-3
    z = w.z
    du, dz = dev(u), dev(z)
    vf = VectorFunction(m)
    R = torch.empty_like(du)
    K = vf.gradient_u(du, dz, residual_out=R)
    close(host(R), o.residual(u, z), RTOL, "residual")
    close(host(K.entries), o.jacobian(u, z, rm, ci), RTOL, "jacobian")
    v, _, gz = o.energy(u, z, False, True)
    sf = PhysicsScalarFunction(m)
    assert abs(sf.value(du, dz) - v) <= RTOL * abs(v)
    close(host(sf.gradient_z(du, dz)), gz, RTOL, "energy grad_z")


@pytest.mark.parametrize("mesh", ["delaunay", "box_general"])
def test_gpu_general_jacobian_bit_identical_rerun(mesh, monkeypatch):
    """The chunk kernel accumulates element blocks in LDS as fixed-point integers, so the
    Jacobian and the residual do not depend on the order in which lanes and waves add
    them: two assemblies are bit-identical (and match the oracle at 1e-12)."""
    print(f"case: general Jacobian rerun, {mesh}", flush=True)
    if mesh == "delaunay":
        m_ = _random_delaunay(400, 3)
        m = ElasticModel(m_, 1.0, 0.3, 3.0, 1e-3)
        X = m_.coords
        conn, nn = m_.conn, m_.num_nodes
    else:
        monkeypatch.setenv("PA_BOX_ASSEMBLY", "0")
        w = wl.make("cantilever_128x64x64", tiny=(32, 16, 16))
        m = wl.model(w)
        assert not m.box_assembly
        X, conn, nn = w.mesh.coords, w.mesh.conn, w.mesh.num_nodes
    rng = np.random.default_rng(17)
    u = rng.normal(size=3 * nn) * 1e-3
    z = rng.uniform(0.05, 1.0, nn)
    vf = VectorFunction(m)
    out = []
    for _ in range(2):
        R = torch.empty(3 * nn, dtype=torch.float64, device="cuda")
        K = vf.gradient_u(dev(u), dev(z), residual_out=R)
        out.append((host(K.entries), host(R)))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    rm, ci = orc.graph(conn, nn)
    o = orc.Model(X, conn, 1.0, 0.3, 3.0, 1e-3)
    close(out[0][0], o.jacobian(u, z, rm, ci), RTOL, "K")


def test_gpu_general_mixed_solve_bit_identical(monkeypatch):
    """On a general mesh (box declared but the chunk kernels forced: PA_BOX_ASSEMBLY=0,
    PA_BOX_MF=0, PA_BOX_RES=0) the FP32 operator, its block diagonal and the FP64
    residual sum in fixed point, so a repeated mixed-precision solve takes the same
    iterations to the same bits; the residual kernel alone matches the oracle."""
    print("case: general-mesh mixed solve rerun", flush=True)
    for k in ("PA_BOX_ASSEMBLY", "PA_BOX_MF", "PA_BOX_RES"):
        monkeypatch.setenv(k, "0")
    w = wl.make("cantilever_128x64x64", tiny=(20, 10, 10))
    m = wl.model(w)
    assert not m.box_assembly
    dz = dev(w.z)
    runs = []
    for _ in range(2):
        pb = Problem(m, tolerance=1e-10, check_every=8, precision="mixed")
        u = host(pb.solution(dz))
        runs.append((u, pb.stats.iterations, pb.stats.outer_iterations))
        assert pb.stats.converged and pb.stats.rel_residual <= 1e-10
    print(f"  iterations {runs[0][1]} and {runs[1][1]}", flush=True)
    assert runs[0][1] == runs[1][1] and runs[0][2] == runs[1][2]
    assert np.array_equal(runs[0][0], runs[1][0])
    u64 = host(Problem(m, tolerance=1e-12).solution(dz))
    close(runs[0][0], u64, 1e-7, "mixed vs FP64 solution")
    # the fixed-point residual against the oracle, and twice the same bits
    rng = np.random.default_rng(5)
    u = rng.normal(size=m.num_dofs) * 1e-2
    vf = VectorFunction(m)
    r1, r2 = host(vf.value(dev(u), dz)), host(vf.value(dev(u), dz))
    assert np.array_equal(r1, r2)
    o = orc.Model(w.mesh.coords, w.mesh.conn, w.cfg.E, w.cfg.nu, w.cfg.p, w.cfg.rho_min,
                  m.tractions[0].face_verts, w.traction)
    close(r1, o.residual(u, w.z), RTOL, "residual")


# ------------------------------------------------------------------ mixed precision
def test_gpu_mixed_precision_solve_cfg0():
    """FP32 matrix-free PCG with FP64 residual replacement reaches the FP64 tolerance on
    the true residual; errors against the FP64 block-CRS PCG solution are bounded."""
    print("case: cfg0 mixed precision", flush=True)
    w = wl.make("cantilever_128x64x64", tiny=(20, 10, 10))
    m = wl.model(w)
    dz = dev(w.z)
    ref = Problem(m, tolerance=1e-12)
    u64 = host(ref.solution(dz))
    g64, C64 = ref.criterionGradient(dz, dev(u64))
    g64 = host(g64)
    mp = Problem(m, tolerance=1e-10, precision="mixed", inner_tolerance=0.1)
    u = mp.solution(dz)
    assert mp.stats.converged and mp.stats.rel_residual <= 1e-10
    assert mp.stats.outer_iterations >= 2  # the residual was actually replaced
    uh = host(u)
    err_u = np.abs(uh - u64).max() / np.abs(u64).max()
    g, C = mp.criterionGradient(dz, u)
    err_c = abs(C - C64) / abs(C64)
    err_g = np.abs(host(g) - g64).max() / np.abs(g64).max()
    print(f"mixed vs fp64: u {err_u:.2e} compliance {err_c:.2e} dC/dz {err_g:.2e} "
          f"outer {mp.stats.outer_iterations} inner {mp.stats.iterations}", flush=True)
    assert err_u < 1e-7 and err_c < 1e-9 and err_g < 1e-7
    # warm start from the converged state: already converged, no inner iteration
    mp2 = Problem(m, tolerance=1e-10, precision="mixed")
    mp2.solution(dz, u0=u)
    assert mp2.stats.iterations == 0 and mp2.stats.converged


@pytest.mark.parametrize("preconditioner", ["jacobi", "multigrid"])
def test_gpu_fp64_warm_start(preconditioner):
    """FP64 solves honour an initial guess: from the converged state no iteration is
    needed; from a perturbed state (constrained dofs included) the solve reaches the same
    solution with its prescribed values restored."""
    print(f"case: fp64 warm start {preconditioner}", flush=True)
    mesh = build_box(8, 4, 4, 2.0, 1.0, 1.0)
    fixed = np.flatnonzero(mesh.coords[:, 0] == 0.0).astype(np.int32)
    bcs = [EssentialBC(fixed, 0, 0.0), EssentialBC(fixed, 1, 0.01), EssentialBC(fixed, 2, -0.02)]
    from platoanalyze_amd.mesh import faces_on_plane
    fv, fe = faces_on_plane(mesh, 0, 2.0)
    m = ElasticModel(mesh, 1.0, 0.3, 3.0, 1e-3, bcs, [Traction(fv, fe, (0.0, 0.0, -1.0))], box=(8, 4, 4))
    z = dev(np.random.default_rng(5).uniform(0.2, 1.0, mesh.num_nodes))
    pb = Problem(m, tolerance=1e-11, preconditioner=preconditioner)
    u = pb.solution(z).clone()
    cold = pb.stats.iterations
    assert pb.stats.converged and cold > 0
    # a looser tolerance than the one u was solved to: its true residual is below it
    loose = Problem(m, tolerance=1e-9, preconditioner=preconditioner)
    loose.solution(z, u0=u)
    assert loose.stats.converged and loose.stats.iterations == 0
    rng = np.random.default_rng(6)
    u0 = u + dev(1e-3 * rng.normal(size=m.num_dofs))
    uw = host(pb.solution(z, u0=u0))
    assert pb.stats.converged and pb.stats.iterations > 0
    assert np.abs(uw[3 * fixed + 1] - 0.01).max() <= 1e-10
    assert np.abs(uw[3 * fixed + 2] + 0.02).max() <= 1e-10
    uc = host(u)
    # both solutions have relative residuals below 1e-11; on this 675-dof problem they
    # agree far inside 1e-6
    err = np.abs(uw - uc).max() / np.abs(uc).max()
    print(f"cold {cold} it, perturbed warm start {pb.stats.iterations} it, difference {err:.2e}", flush=True)
    assert err <= 1e-6


def test_gpu_mixed_precision_nonzero_dirichlet():
    mesh = build_box(8, 4, 4, 2.0, 1.0, 1.0)
    fixed = np.flatnonzero(mesh.coords[:, 0] == 0.0).astype(np.int32)
    bcs = [EssentialBC(fixed, 0, 0.0), EssentialBC(fixed, 1, 0.01), EssentialBC(fixed, 2, -0.02)]
    from platoanalyze_amd.mesh import faces_on_plane
    fv, fe = faces_on_plane(mesh, 0, 2.0)
    m = ElasticModel(mesh, 1.0, 0.3, 3.0, 1e-3, bcs, [Traction(fv, fe, (0.0, 0.0, -1.0))])
    z = dev(np.random.default_rng(9).uniform(0.2, 1.0, mesh.num_nodes))
    u64 = host(Problem(m, tolerance=1e-12).solution(z))
    mp = Problem(m, tolerance=1e-10, precision="mixed")
    uh = host(mp.solution(z))
    assert mp.stats.converged
    np.testing.assert_allclose(uh[3 * fixed + 1], 0.01, rtol=0, atol=0)
    assert np.abs(uh - u64).max() <= 1e-7 * np.abs(u64).max()


# ------------------------------------------------------------------ stress p-norm + adjoint
def test_gpu_stress_pnorm_goldens_2x2x2(box222, goldens):
    from platoanalyze_amd.criteria import StressPNorm

    m = ElasticModel(box222, 1e6, 0.3, 1.0, 0.0)
    u, z = dev(ref_state(27)), dev(np.ones(27))
    g = goldens["stress_pnorm3d"]
    sp = StressPNorm(m, 12.0, "voigt")
    assert abs(sp.value(u, z) - g["value_gold"]) <= 1e-13 * g["value_gold"]
    close(host(sp.gradient_u(u, z))[:27], g["grad_u_gold"], 1e-12, "pnorm grad_u")
    close(host(sp.gradient_z(u, z)), g["grad_z_gold"], 1e-12, "pnorm grad_z")


@pytest.mark.parametrize("measure,pn,q", [("voigt", 8.0, 0.0), ("von_mises", 8.0, 0.5), ("von_mises", 3.5, 1.0)])
def test_gpu_stress_pnorm_partials_cfg0(cfg0, measure, pn, q):
    from platoanalyze_amd.criteria import StressPNorm

    w, m, o, rm, ci, u, z = cfg0
    sp = StressPNorm(m, pn, measure, q)
    du, dz = dev(u), dev(z)
    F, gu, gz = o.pnorm(u, z, 0 if measure == "voigt" else 1, pn, q)
    assert abs(sp.value(du, dz) - F) <= RTOL * abs(F)
    close(host(sp.gradient_u(du, dz)), gu, RTOL, "dF/du")
    close(host(sp.gradient_z(du, dz)), gz, RTOL, "dF/dz partial")


def test_gpu_dRdz_T_apply_cfg0(cfg0):
    from platoanalyze_amd.criteria import dRdz_T_apply

    w, m, o, rm, ci, u, z = cfg0
    lam = np.random.default_rng(31).normal(size=u.size)
    close(host(dRdz_T_apply(m, dev(u), dev(lam), dev(z))), o.dRdz_T_apply(u, lam, z), RTOL, "(dR/dz)^T lambda")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_pnorm_total_gradient_fd_and_autograd(precision):
    """Hand-written adjoint of the relaxed von Mises p-norm (p=8, q=0.5) through a
    torch.autograd.Function, with forward and adjoint solves in the precision under test,
    against central finite differences (h=1e-6) of full FP64 solves: the FD reference
    needs F to ~1e-13, which only the FP64 solve delivers without FP32 rounding noise."""
    from platoanalyze_amd import criteria

    w = wl.make("cantilever_20x10x10", tiny=(8, 4, 4))
    m = wl.model(w)
    z0 = np.random.default_rng(13).uniform(0.3, 1.0, w.mesh.num_nodes)
    pb = Problem(m, tolerance=1e-12, precision=precision, inner_tolerance=0.1)
    ref = Problem(m, tolerance=1e-13)
    z = dev(z0).requires_grad_(True)
    F = criteria.stress_pnorm(z, pb, 8.0, "von_mises", 0.5)
    F.backward()
    g = host(z.grad)
    C = criteria.compliance(z.detach(), pb)
    rng = np.random.default_rng(64)
    for n in rng.choice(w.mesh.num_nodes, 6, replace=False):
        h = 1e-6
        zp, zm = z0.copy(), z0.copy()
        zp[n] += h
        zm[n] -= h
        fp = float(criteria.stress_pnorm(dev(zp), ref, 8.0, "von_mises", 0.5))
        fm = float(criteria.stress_pnorm(dev(zm), ref, 8.0, "von_mises", 0.5))
        fd = (fp - fm) / (2 * h)
        print(f"node {n}: FD {fd:.10e} adjoint {g[n]:.10e}", flush=True)
        assert abs(fd - g[n]) <= 1e-6 * np.abs(g).max() + 1e-7 * abs(fd), (n, fd, g[n])
    assert float(C) > 0


# ------------------------------------------------------------------ shape derivatives
def test_gpu_gradient_x_goldens_2x2x2(box222, goldens):
    """VectorFunction::gradient_x (DerivativeTests.cpp:469-486, 36 entries, transposed
    blocks in the reference's row-0 order) and PhysicsScalarFunction::gradient_x of the
    internal elastic energy (:660-693, 81 values)."""
    print("case: 2x2x2 gradient_x goldens", flush=True)
    rm, ci = orc.graph(box222.conn, 27)
    ci = ci.copy()
    ci[rm[0]:rm[1]] = REF_ROW0
    m = ElasticModel(box222, 1e6, 0.3, 1.0, 0.0, graph=(rm, ci))
    u, z = dev(ref_state(27)), dev(np.ones(27))
    G = host(VectorFunction(m).gradient_x(u, z).entries)
    gold = np.array(goldens["elastostatic_residual3d"]["gold_grad_x_entries"])
    np.testing.assert_allclose(G[:36], gold, rtol=1e-13, atol=1e-9)
    gx = host(PhysicsScalarFunction(m).gradient_x(u, z))
    close(gx, goldens["internal_elastic_energy3d"]["grad_x_gold"], 1e-12, "energy grad_x")


def test_gpu_gradient_x_cfg0(cfg0):
    """dR/dX (with the traction's face dependence) and d(energy)/dX against the oracle's
    dual-number evaluation on config 0."""
    print("case: cfg0 gradient_x", flush=True)
    w, m, o, rm, ci, u, z = cfg0
    du, dz = dev(u), dev(z)
    close(host(VectorFunction(m).gradient_x(du, dz).entries), o.gradient_x(u, z, rm, ci), RTOL, "dR/dX")
    close(host(PhysicsScalarFunction(m).gradient_x(du, dz)), o.energy_gradient_x(u, z), RTOL, "dE/dX")


@pytest.mark.parametrize("seed", [0])
def test_gpu_gradient_x_unstructured(seed):
    """Unstructured tets with rows wider than 16 blocks (8-bit slot codes)."""
    mesh = _random_delaunay(400, seed)
    rm, ci = orc.graph(mesh.conn, mesh.num_nodes)
    m = ElasticModel(mesh, 2.0, 0.25, 3.0, 1e-3)
    o = orc.Model(mesh.coords, mesh.conn, 2.0, 0.25, 3.0, 1e-3)
    rng = np.random.default_rng(seed + 20)
    u = rng.normal(size=3 * mesh.num_nodes)
    z = rng.uniform(0.05, 1.0, mesh.num_nodes)
    du, dz = dev(u), dev(z)
    close(host(VectorFunction(m).gradient_x(du, dz).entries), o.gradient_x(u, z, rm, ci), RTOL, "dR/dX")
    close(host(PhysicsScalarFunction(m).gradient_x(du, dz)), o.energy_gradient_x(u, z), RTOL, "dE/dX")


# ------------------------------------------------------------------ penalty models, volume
PENALTIES = [("SIMP", 0, 3.0, 1e-3, 1.0), ("RAMP", 1, 3.0, 1e-3, 1.0), ("Heaviside", 2, 2.0, 1e-3, 1.2),
             ("NoPenalty", 3, 1.0, 0.0, 1.0)]


@pytest.mark.parametrize("name,kind,p,rmin,L", PENALTIES)
def test_gpu_penalty_models_cfg0(cfg0, name, kind, p, rmin, L):
    """Residual, Jacobian, dR/dz, energy and dC/dz under each "Penalty Function" against
    the oracle; geometric Volume (value, dV/dz, dV/dX) with the same penalty."""
    from platoanalyze_amd.elliptic import Volume

    print(f"case: cfg0 penalty {name}", flush=True)
    w, _, _, rm, ci, u, z = cfg0
    bcs = [EssentialBC(w.fixed_nodes, d, 0.0) for d in range(3)]
    m = ElasticModel(w.mesh, w.cfg.E, w.cfg.nu, p, rmin, bcs, [Traction(w.face_verts, w.face_elems, w.traction)],
                     penalty_model=name, regularization_length=L)
    o = orc.Model(w.mesh.coords, w.mesh.conn, w.cfg.E, w.cfg.nu, p, rmin, w.face_verts, w.traction,
                  penalty_model=kind, length=L)
    du, dz = dev(u), dev(z)
    vf, sf = VectorFunction(m), PhysicsScalarFunction(m)
    R = torch.empty_like(du)
    K = vf.gradient_u(du, dz, residual_out=R)
    close(host(R), o.residual(u, z), RTOL, "residual")
    close(host(K.entries), o.jacobian(u, z, rm, ci), RTOL, "jacobian")
    close(host(vf.gradient_z(du, dz).entries), o.gradient_z(u, z, rm, ci), RTOL, "dR/dz")
    v, _, gz = o.energy(u, z, False, True)
    assert abs(sf.value(du, dz) - v) <= RTOL * abs(v)
    if kind != 3:
        close(host(sf.gradient_z(du, dz)), gz, RTOL, "dC/dz")
    else:
        assert np.abs(host(sf.gradient_z(du, dz))).max() == 0.0
    vol = Volume(m, name, p, rmin, L)
    ov, ogz, ogx = o.volume(z, kind, p, rmin, L)
    assert abs(vol.value(dz) - ov) <= RTOL * abs(ov)
    if kind != 3:
        close(host(vol.gradient_z(dz)), ogz, RTOL, "dV/dz")
    close(host(vol.gradient_x(dz)), ogx, RTOL, "dV/dX")


@pytest.mark.parametrize("name,p,L", [("RAMP", 3.0, 1.0), ("Heaviside", 2.0, 1.2)])
def test_gpu_penalty_models_solves(name, p, L):
    """Mixed-precision (FP32 matrix-free operator) and multigrid solves under non-SIMP
    penalties agree with the FP64 Jacobi solve."""
    mesh = build_box(16, 8, 8, 2.0, 1.0, 1.0)
    from platoanalyze_amd.mesh import faces_on_plane
    fixed = np.flatnonzero(mesh.coords[:, 0] == 0.0).astype(np.int32)
    fv, fe = faces_on_plane(mesh, 0, 2.0)
    m = ElasticModel(mesh, 1.0, 0.3, p, 1e-3, [EssentialBC(fixed, d, 0.0) for d in range(3)],
                     [Traction(fv, fe, (0.0, 0.0, -1.0))], box=(16, 8, 8), penalty_model=name,
                     regularization_length=L)
    z = dev(np.random.default_rng(4).uniform(0.2, 1.0, mesh.num_nodes))
    u64 = host(Problem(m, tolerance=1e-12).solution(z))
    for prec, pre in (("mixed", "jacobi"), ("mixed", "multigrid"), ("fp64", "multigrid")):
        pb = Problem(m, tolerance=1e-10, precision=prec, preconditioner=pre)
        u = host(pb.solution(z))
        err = np.abs(u - u64).max() / np.abs(u64).max()
        print(f"{name} {prec}/{pre}: {pb.stats.iterations} it, err {err:.1e}", flush=True)
        assert pb.stats.converged and err <= 1e-7


def test_gpu_model_from_exodus(tmp_path):
    """A problem built from an Exodus file's named sets (node set -> essential BCs, side
    set -> traction) gives the residual and solution of the geometric workload."""
    from platoanalyze_amd.exodus import model_from_exodus, read_exodus, sides_of_faces, write_exodus

    w = wl.make("cantilever_20x10x10", tiny=(8, 4, 4))
    el, sd = sides_of_faces(w.mesh, w.face_verts, w.face_elems)
    path = tmp_path / "cantilever.exo"
    write_exodus(path, w.mesh, {"x-": w.fixed_nodes}, {"x+": (el, sd)})
    ex = read_exodus(path)
    m = model_from_exodus(ex, w.cfg.E, w.cfg.nu, w.cfg.p, w.cfg.rho_min,
                          essential={"x-": [(0, 0.0), (1, 0.0), (2, 0.0)]}, tractions={"x+": w.traction})
    ref = wl.model(w)
    z = dev(w.z)
    u = dev(np.random.default_rng(2).normal(size=3 * w.mesh.num_nodes))
    close(host(VectorFunction(m).value(u, z)), host(VectorFunction(ref).value(u, z)), RTOL, "residual")
    close(host(Problem(m, tolerance=1e-12).solution(z)), host(Problem(ref, tolerance=1e-12).solution(z)), 1e-10, "u")
    # the file's box lattice is declared: lattice assembly and multigrid apply
    import ctypes as C

    from platoanalyze_amd import _capi
    v = C.c_int(-1)
    _capi.call("pa_box_assembly", m.handle, C.byref(v))
    assert v.value == 1
    close(host(Problem(m, tolerance=1e-10, preconditioner="multigrid").solution(z)),
          host(Problem(ref, tolerance=1e-12).solution(z)), 1e-7, "u (multigrid)")
    # a graded lattice is no box: the general assembly, the same residual as a direct model
    X = w.mesh.coords.copy()
    X[:, 0] = X[:, 0] ** 1.3 / 2.0 ** 0.3
    ex.mesh = TetMesh(X, w.mesh.conn)
    mg = model_from_exodus(ex, w.cfg.E, w.cfg.nu, w.cfg.p, w.cfg.rho_min,
                           essential={"x-": [(0, 0.0), (1, 0.0), (2, 0.0)]}, tractions={"x+": w.traction})
    _capi.call("pa_box_assembly", mg.handle, C.byref(v))
    assert v.value == 0
