"""The oracle (CPU restatement) pinned against the reference's own golden vectors and
against closed forms / invariants.  CPU only.

Reference test inputs restated here (unit_tests/DerivativeTests.cpp):
  mesh     PlatoUtestHelpers::getBoxMesh(3, 2) = Omega_h::build_box 1x1x1, 2x2x2 hexes
  u        u[i] = (disp += 1e-4), i.e. accumulated, i = 0..3N-1  (:294-299)
  z        1.0 everywhere
  material Isotropic Linear Elastic, E = 1e6, nu = 0.3
  penalty  SIMP Exponent 1, Minimum Value 0
The reference's "first row" CRS column order of the 2x2x2 box (Omega_h edge order) is
pinned by the Jacobian / gradient_z goldens: 0, 8, 1, 3, 9, 13, 21, 25.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from platoanalyze_amd.mesh import build_box, faces_on_plane, nodes_on_plane

REF_ROW0 = [0, 8, 1, 3, 9, 13, 21, 25]


def ref_state(nn):
    u = np.empty(3 * nn)
    d = 0.0
    for i in range(3 * nn):
        d += 1e-4
        u[i] = d
    return u


def assert_close(got, gold, rtol):
    """TEST_FLOATING_EQUALITY per entry; golden zeros checked against the vector norm."""
    got = np.asarray(got, float)
    gold = np.asarray(gold, float)
    assert got.shape == gold.shape
    scale = np.abs(gold).max()
    for g, r in zip(got, gold):
        if r == 0.0:
            assert abs(g) <= 1e-12 * max(scale, 1.0), (g, r)
        else:
            assert abs(g - r) <= rtol * abs(r), (g, r, abs(g - r) / abs(r))


def reference_graph(conn, nn):
    """Sorted-star graph with row 0 in the reference's Omega_h edge order."""
    rm, ci = orc.graph(conn, nn)
    ci = ci.copy()
    assert sorted(ci[rm[0]:rm[1]]) == sorted(REF_ROW0)
    ci[rm[0]:rm[1]] = REF_ROW0
    return rm, ci


@pytest.fixture(scope="module")
def ref_model(box222):
    return orc.Model(box222.coords, box222.conn, E=1e6, nu=0.3, p=1.0, rmin=0.0)


def test_box_1x1x1_matches_reference_element_coords(goldens):
    m = build_box(1, 1, 1)
    gold = np.array(goldens["local_element_coords_1x1x1"])
    got = m.coords[m.conn]
    np.testing.assert_array_equal(got, gold)


def test_box_2x2x2_sizes(box222):
    assert box222.num_nodes == 27 and box222.num_elems == 48
    # all tets positive volume magnitude 1/48 of the unit box
    X = box222.coords[box222.conn]
    J = np.stack([X[:, 0] - X[:, 3], X[:, 1] - X[:, 3], X[:, 2] - X[:, 3]], axis=2)
    vol = np.abs(np.linalg.det(J)) / 6
    np.testing.assert_allclose(vol, 1.0 / 48, rtol=1e-14)


def test_residual_golden(ref_model, box222, goldens):
    """ElastostaticResidual3D value (DerivativeTests.cpp:361-398); pins node numbering."""
    u = ref_state(27)
    R = ref_model.residual(u, np.ones(27))
    assert_close(R, goldens["elastostatic_residual3d"]["residual_gold"], 1e-13)


def test_cell_worksets_golden(ref_model, box222, goldens):
    """DerivativeTests "3D" (:159-253): the six golden cells are found in this mesh by
    their gradients+strain, and all four worksets match to 1e-13."""
    g = goldens["cells3d"]
    gg = np.array(g["gradient_gold"]).reshape(6, 4, 3)
    sg = np.array(g["strain_gold"]).reshape(6, 6)
    tg = np.array(g["stress_gold"]).reshape(6, 6)
    rg = np.array(g["result_gold"]).reshape(6, 12)
    u = ref_state(27)
    grad, vol, eps, sig, res = ref_model.cell_worksets(u, 48)
    for k in range(6):
        hits = [c for c in range(48)
                if np.allclose(grad[c], gg[k], rtol=1e-13, atol=1e-12)
                and np.allclose(eps[c], sg[k], rtol=1e-13, atol=1e-15)]
        assert hits, f"golden cell {k} not found"
        c = hits[0]
        assert_close(grad[c].ravel(), gg[k].ravel(), 1e-13)
        assert_close(eps[c], sg[k], 1e-13)
        assert_close(sig[c], tg[k], 1e-13)
        assert_close(res[c], rg[k], 1e-13)


def test_jacobian_golden(ref_model, box222, goldens):
    """ElastostaticResidual3D gradient_u (:408-428), first 45 entries, tol 1e-15 as the reference."""
    rm, ci = reference_graph(box222.conn, 27)
    K = ref_model.jacobian(ref_state(27), np.ones(27), rm, ci)
    gold = np.array(goldens["elastostatic_residual3d"]["gold_jac_entries"])
    np.testing.assert_allclose(K[:45], gold, rtol=1e-15, atol=1e-10)


def test_gradient_z_golden(ref_model, box222, goldens):
    """ElastostaticResidual3D gradient_z (:444-453), transposed 1x3 blocks, tol 2e-14."""
    rm, ci = reference_graph(box222.conn, 27)
    G = ref_model.gradient_z(ref_state(27), np.ones(27), rm, ci)
    assert_close(G[:24], goldens["elastostatic_residual3d"]["gold_grad_entries"], 2e-14)


def test_energy_golden(ref_model, goldens):
    """InternalElasticEnergy3D value / grad_u / grad_z (:585-652)."""
    g = goldens["internal_elastic_energy3d"]
    v, gu, gz = ref_model.energy(ref_state(27), np.ones(27))
    assert abs(v - g["value_gold"]) <= 1e-13 * g["value_gold"]
    assert_close(gu, g["grad_u_gold"], 1e-12)  # gold printed with 15 significant digits
    assert_close(gz, g["grad_z_gold"], 1e-13)


def test_gradient_x_golden(ref_model, box222, goldens):
    """ElastostaticResidual3D gradient_x (:469-486): transposed 3x3 blocks, first 36
    entries, tol 1e-13 as the reference."""
    rm, ci = reference_graph(box222.conn, 27)
    G = ref_model.gradient_x(ref_state(27), np.ones(27), rm, ci)
    assert_close(G[:36], goldens["elastostatic_residual3d"]["gold_grad_x_entries"], 1e-13)


def test_energy_gradient_x_golden(ref_model, goldens):
    """InternalElasticEnergy3D gradient_x (:660-693), 81 values, tol 1e-13 as the reference
    (gold printed with 15 significant digits: 1e-12 where it is short)."""
    g = ref_model.energy_gradient_x(ref_state(27), np.ones(27))
    assert_close(g, goldens["internal_elastic_energy3d"]["grad_x_gold"], 1e-12)


def test_stress_pnorm_golden(ref_model, goldens):
    """StressPNorm3D value (:1058), exponent 12, and its partial gradients (:1069-1108)."""
    g = goldens["stress_pnorm3d"]
    v = ref_model.stress_pnorm(ref_state(27), np.ones(27), 12.0)
    assert abs(v - g["value_gold"]) <= 1e-13 * g["value_gold"]
    v2, gu, gz = ref_model.pnorm(ref_state(27), np.ones(27), 0, 12.0)
    assert v2 == v
    assert_close(gu[:27], g["grad_u_gold"], 1e-12)
    assert_close(gz, g["grad_z_gold"], 1e-12)


# ----------------------------------------------------------------- invariants
@pytest.fixture(scope="module")
def small():
    m = build_box(4, 3, 2, 2.0, 1.0, 1.0)
    rng = np.random.default_rng(7)
    z = rng.uniform(0.2, 1.0, m.num_nodes)
    u = rng.normal(size=3 * m.num_nodes) * 1e-2
    model = orc.Model(m.coords, m.conn, E=1.0, nu=0.3, p=3.0, rmin=1e-3)
    rm, ci = orc.graph(m.conn, m.num_nodes)
    return m, z, u, model, rm, ci


def test_jacobian_symmetric_rigid_modes_fd(small):
    m, z, u, model, rm, ci = small
    K = orc.to_dense(rm, ci, model.jacobian(u, z, rm, ci))
    np.testing.assert_allclose(K, K.T, atol=1e-13 * np.abs(K).max())
    X = m.coords
    modes = [np.tile(e, m.num_nodes) for e in np.eye(3)]
    for axis in range(3):  # infinitesimal rotations
        w = np.zeros(3)
        w[axis] = 1.0
        modes.append(np.cross(w, X).ravel())
    for v in modes:
        assert np.abs(K @ v).max() <= 1e-12 * np.abs(K).max() * np.abs(v).max()
    # K = dR/du: residual is linear in u, so a central difference is exact up to rounding
    d = np.random.default_rng(3).normal(size=u.size)
    h = 1e-3
    fd = (model.residual(u + h * d, z) - model.residual(u - h * d, z)) / (2 * h)
    np.testing.assert_allclose(K @ d, fd, atol=1e-9 * np.abs(fd).max())
    # 6 rigid modes => exactly 6 zero eigenvalues, the rest positive
    ev = np.linalg.eigvalsh(K)
    assert (np.abs(ev[:6]) < 1e-10 * ev[-1]).all() and ev[6] > 1e-8 * ev[-1]


def test_residual_is_K_u_minus_f_and_patch(small):
    m, z, u, model, rm, ci = small
    fv, fe = faces_on_plane(m, 0, 2.0)
    mt = orc.Model(m.coords, m.conn, E=1.0, nu=0.3, p=3.0, rmin=1e-3, face_verts=fv, traction=(0.0, 0.0, -1.0))
    K = orc.to_dense(rm, ci, model.jacobian(u, z, rm, ci))
    R = mt.residual(u, z)
    f = -mt.residual(np.zeros_like(u), z)
    np.testing.assert_allclose(R, K @ u - f, atol=1e-12 * np.abs(R).max())
    # traction: total force = face area (1x1) * t
    F = f.reshape(-1, 3).sum(0)
    np.testing.assert_allclose(F, [0, 0, -1.0], atol=1e-14)
    # patch test: affine displacement => zero residual at interior nodes
    A = np.array([[1e-3, 2e-3, -1e-3], [0.5e-3, -1e-3, 3e-3], [2e-3, 1e-3, 1e-3]])
    ua = (m.coords @ A.T).ravel()
    Ra = model.residual(ua, np.full_like(z, 0.7)).reshape(-1, 3)  # uniform material
    X = m.coords
    interior = np.all((X > 1e-9) & (X < np.array([2.0, 1.0, 1.0]) - 1e-9), axis=1)
    assert interior.any()
    assert np.abs(Ra[interior]).max() < 1e-14


def test_energy_closed_forms(small):
    m, z, u, model, rm, ci = small
    K = orc.to_dense(rm, ci, model.jacobian(u, z, rm, ci))
    v, gu, gz = model.energy(u, z)
    np.testing.assert_allclose(v, 0.5 * u @ K @ u, rtol=1e-13)
    np.testing.assert_allclose(gu, K @ u, atol=1e-13 * np.abs(gu).max())
    # dC/dz by central differences on a few nodes
    for n in [0, 5, 17, m.num_nodes - 1]:
        h = 1e-6
        zp, zm = z.copy(), z.copy()
        zp[n] += h
        zm[n] -= h
        fd = (model.energy(u, zp, False, False)[0] - model.energy(u, zm, False, False)[0]) / (2 * h)
        assert abs(fd - gz[n]) <= 1e-6 * abs(gz[n]) + 1e-12


def test_gradient_z_fd(small):
    m, z, u, model, rm, ci = small
    G = model.gradient_z(u, z, rm, ci)  # rows = control node, 1x3 blocks
    for n in [1, 9, m.num_nodes // 2]:
        h = 1e-6
        zp, zm = z.copy(), z.copy()
        zp[n] += h
        zm[n] -= h
        fd = (model.residual(u, zp) - model.residual(u, zm)) / (2 * h)
        col = np.zeros(3 * m.num_nodes)
        for s in range(rm[n], rm[n + 1]):
            col[3 * ci[s]:3 * ci[s] + 3] = G[3 * s:3 * s + 3]
        np.testing.assert_allclose(col, fd, atol=1e-6 * np.abs(fd).max())


def test_gradient_x_fd_with_traction(small):
    """dR/dX (transposed blocks, including the traction's dependence on the loaded faces)
    and d(energy)/dX against central differences of the residual and energy in the nodal
    coordinates; the loaded nodes are among the probed ones."""
    m, z, u, _, rm, ci = small
    fv, fe = faces_on_plane(m, 0, 2.0)
    t = np.array([0.3, -0.2, -1.0])
    model = orc.Model(m.coords, m.conn, E=1.0, nu=0.3, p=3.0, rmin=1e-3, face_verts=fv, traction=t)
    G = model.gradient_x(u, z, rm, ci)
    gE = model.energy_gradient_x(u, z)
    probe = [1, m.num_nodes // 2, int(fv[0, 0]), int(fv[-1, 1])]
    for n in probe:
        for k in range(3):
            h = 1e-6
            Xp, Xm = m.coords.copy(), m.coords.copy()
            Xp[n, k] += h
            Xm[n, k] -= h
            mp = orc.Model(Xp, m.conn, E=1.0, nu=0.3, p=3.0, rmin=1e-3, face_verts=fv, traction=t)
            mm = orc.Model(Xm, m.conn, E=1.0, nu=0.3, p=3.0, rmin=1e-3, face_verts=fv, traction=t)
            fd = (mp.residual(u, z) - mm.residual(u, z)) / (2 * h)
            col = np.zeros(3 * m.num_nodes)
            for s in range(rm[n], rm[n + 1]):
                col[3 * ci[s]:3 * ci[s] + 3] = G[9 * s + 3 * k:9 * s + 3 * k + 3]
            np.testing.assert_allclose(col, fd, atol=1e-6 * np.abs(fd).max() + 1e-12)
            fe_ = (mp.energy(u, z, False, False)[0] - mm.energy(u, z, False, False)[0]) / (2 * h)
            assert abs(fe_ - gE[3 * n + k]) <= 1e-6 * np.abs(gE).max(), (n, k, fe_, gE[3 * n + k])


def test_constraints_and_pcg_vs_dense(small):
    m, z, u, model, rm, ci = small
    fv, fe = faces_on_plane(m, 0, 2.0)
    mt = orc.Model(m.coords, m.conn, E=1.0, nu=0.3, p=3.0, rmin=1e-3, face_verts=fv, traction=(0.0, 0.0, -1.0))
    fixed = nodes_on_plane(m, 0, 0.0)
    dofs = np.sort(np.concatenate([3 * fixed + d for d in range(3)]))
    f = -mt.residual(np.zeros(3 * m.num_nodes), z)
    K = model.jacobian(None, z, rm, ci)
    Kc, bc = orc.apply_constraints(rm, ci, K, f, dofs, np.zeros(len(dofs)))
    Kd = orc.to_dense(rm, ci, Kc)
    free = np.setdiff1d(np.arange(3 * m.num_nodes), dofs)
    # constrained rows/cols: identity; free block untouched
    np.testing.assert_array_equal(Kd[np.ix_(dofs, dofs)], np.eye(len(dofs)))
    assert np.all(Kd[np.ix_(dofs, free)] == 0) and np.all(Kd[np.ix_(free, dofs)] == 0)
    Kfull = orc.to_dense(rm, ci, K)
    np.testing.assert_array_equal(Kd[np.ix_(free, free)], Kfull[np.ix_(free, free)])
    x_dense = np.linalg.solve(Kd, bc)
    x, it, rel = orc.pcg(rm, ci, Kc, bc, tol=1e-12)
    assert rel < 1e-11
    np.testing.assert_allclose(x, x_dense, atol=1e-9 * np.abs(x_dense).max())
    # compliance at equilibrium: C = 0.5 f.u
    v, _, _ = model.energy(x, z, False, False)
    np.testing.assert_allclose(v, 0.5 * f @ x, rtol=1e-9)


def test_nonzero_dirichlet_lift(small):
    """applyBlockConstraints with prescribed values: solution carries the values and the
    free part solves K_ff x_f = b_f - K_fc x_c."""
    m, z, u, model, rm, ci = small
    K = model.jacobian(None, z, rm, ci)
    fixed = nodes_on_plane(m, 0, 0.0)
    dofs = np.concatenate([3 * fixed, 3 * fixed + 1, 3 * fixed + 2])  # removes every rigid mode
    vals = np.concatenate([np.zeros(len(fixed)), np.linspace(0.01, 0.02, len(fixed)), np.zeros(len(fixed))])
    b = np.random.default_rng(1).normal(size=3 * m.num_nodes)
    Kc, bc = orc.apply_constraints(rm, ci, K, b, dofs, vals, scale=1.0)
    x = np.linalg.solve(orc.to_dense(rm, ci, Kc), bc)
    np.testing.assert_allclose(x[dofs], vals, rtol=1e-12)
    Kf = orc.to_dense(rm, ci, K)
    free = np.setdiff1d(np.arange(3 * m.num_nodes), dofs)
    lhs = Kf[np.ix_(free, free)] @ x[free]
    rhs = b[free] - Kf[np.ix_(free, dofs)] @ vals
    np.testing.assert_allclose(lhs, rhs, atol=1e-10 * np.abs(rhs).max())


def test_pnorm_von_mises_fd(small):
    """Relaxed von Mises p-norm (no reference counterpart): partials against central
    differences, and the homogeneity dF/dz = q F / (4 rho) per element for uniform rho."""
    m, z, u, model, rm, ci = small
    for measure, pn, q in [(1, 8.0, 0.5), (0, 8.0, 0.0)]:
        F, gu, gz = model.pnorm(u, z, measure, pn, q)
        d = np.random.default_rng(8).normal(size=u.size)
        h = 1e-6
        fd = (model.pnorm(u + h * d, z, measure, pn, q, False, False)[0]
              - model.pnorm(u - h * d, z, measure, pn, q, False, False)[0]) / (2 * h)
        assert abs(fd - gu @ d) <= 1e-6 * abs(fd)
        for n in [0, 7, m.num_nodes // 2]:
            zp, zm = z.copy(), z.copy()
            zp[n] += h
            zm[n] -= h
            fdz = (model.pnorm(u, zp, measure, pn, q, False, False)[0]
                   - model.pnorm(u, zm, measure, pn, q, False, False)[0]) / (2 * h)
            assert abs(fdz - gz[n]) <= 1e-6 * np.abs(gz).max()
    # uniform density: F(z) = rho^q F0  => sum_n dF/dz_n = q F / rho
    zc = np.full_like(z, 0.4)
    F, _, gz = model.pnorm(u, zc, 1, 8.0, 0.5, False, True)
    np.testing.assert_allclose(gz.sum(), 0.5 * F / 0.4, rtol=1e-12)


def test_dRdz_T_apply_matches_gradient_z(small):
    m, z, u, model, rm, ci = small
    lam = np.random.default_rng(12).normal(size=u.size)
    G = model.gradient_z(u, z, rm, ci)  # rows = control node, entries dR_(col,i)/dz_row
    want = np.zeros(m.num_nodes)
    for r in range(m.num_nodes):
        for s in range(rm[r], rm[r + 1]):
            want[r] += G[3 * s:3 * s + 3] @ lam[3 * ci[s]:3 * ci[s] + 3]
    np.testing.assert_allclose(model.dRdz_T_apply(u, lam, z), want, atol=1e-12 * np.abs(want).max())


# ----------------------------------------------------------------- penalty models, volume
def test_volume_closed_forms(small):
    """V = box volume for NoPenalty; linear SIMP (p = 1, rmin = 0) at uniform z = c gives
    c times it with dV/dz summing to the box volume; translation invariance of dV/dX."""
    m, z, u, model, rm, ci = small
    box = 2.0 * 1.0 * 1.0
    v, gz, gx = model.volume(z, model=3)
    assert abs(v - box) <= 1e-13 * box
    np.testing.assert_array_equal(gz, 0.0)
    np.testing.assert_allclose(gx.reshape(-1, 3).sum(axis=0), 0.0, atol=1e-13)
    v, gz, gx = model.volume(np.full(m.num_nodes, 0.4), model=0, p=1.0, rmin=0.0)
    assert abs(v - 0.4 * box) <= 1e-13 * box
    assert abs(gz.sum() - box) <= 1e-13 * box


@pytest.mark.parametrize("pm", [(0, 3.0, 1e-3, 1.0), (1, 3.0, 1e-3, 1.0), (2, 2.0, 1e-3, 1.2), (3, 1.0, 0.0, 1.0)])
def test_volume_fd(small, pm):
    m, z, u, model, rm, ci = small
    kind, p, rmin, L = pm
    v, gz, gx = model.volume(z, kind, p, rmin, L)
    for n in [2, m.num_nodes // 3]:
        h = 1e-6
        zp, zm = z.copy(), z.copy()
        zp[n] += h
        zm[n] -= h
        fd = (model.volume(zp, kind, p, rmin, L)[0] - model.volume(zm, kind, p, rmin, L)[0]) / (2 * h)
        assert abs(fd - gz[n]) <= 1e-7 * max(np.abs(gz).max(), 1e-30) + 1e-12, (n, fd, gz[n])
        for k in range(3):
            Xp, Xm = m.coords.copy(), m.coords.copy()
            Xp[n, k] += h
            Xm[n, k] -= h
            vp = orc.Model(Xp, m.conn).volume(z, kind, p, rmin, L)[0]
            vm = orc.Model(Xm, m.conn).volume(z, kind, p, rmin, L)[0]
            assert abs((vp - vm) / (2 * h) - gx[3 * n + k]) <= 1e-7 * np.abs(gx).max(), (n, k)


@pytest.mark.parametrize("pm", [(1, 3.0, 1e-3, 1.0), (2, 2.0, 1e-3, 1.2), (3, 1.0, 0.0, 1.0)])
def test_penalty_models_fd_and_limits(small, pm):
    """RAMP, Heaviside and NoPenalty in the residual, Jacobian and energy: dC/dz and dR/dz
    against central differences; the Jacobian against differences of the residual; and
    each model's closed form: RAMP(1) = SIMP(1) = 1, Heaviside saturates at 1 for z >= L,
    NoPenalty equals SIMP with exponent 0."""
    m, z, u, _, rm, ci = small
    kind, p, rmin, L = pm
    mod = orc.Model(m.coords, m.conn, E=1.0, nu=0.3, p=p, rmin=rmin, penalty_model=kind, length=L)
    v, gu, gz = mod.energy(u, z)
    G = mod.gradient_z(u, z, rm, ci)
    K = mod.jacobian(u, z, rm, ci)
    R0 = mod.residual(u, z)
    for n in [1, m.num_nodes // 2]:
        h = 1e-6
        zp, zm = z.copy(), z.copy()
        zp[n] += h
        zm[n] -= h
        fd = (mod.energy(u, zp, False, False)[0] - mod.energy(u, zm, False, False)[0]) / (2 * h)
        # floors: rounding of the differenced values (|value| * 1e-16 / h)
        assert abs(fd - gz[n]) <= 1e-6 * np.abs(gz).max() + 1e-9 * abs(v), (n, fd, gz[n])
        fdr = (mod.residual(u, zp) - mod.residual(u, zm)) / (2 * h)
        col = np.zeros(3 * m.num_nodes)
        for s in range(rm[n], rm[n + 1]):
            col[3 * ci[s]:3 * ci[s] + 3] = G[3 * s:3 * s + 3]
        np.testing.assert_allclose(col, fdr, atol=1e-6 * np.abs(fdr).max() + 1e-9 * np.abs(R0).max())
    # the residual is linear in u: R(u) = K u - f (no traction here)
    Kd = orc.to_dense(rm, ci, K)
    np.testing.assert_allclose(Kd @ u, R0, atol=1e-10 * np.abs(R0).max())
    ones = np.ones(m.num_nodes)
    simp1 = orc.Model(m.coords, m.conn, E=1.0, nu=0.3, p=1.0, rmin=0.0)
    if kind in (1, 3):  # RAMP at z = 1 and NoPenalty anywhere: full stiffness
        zz = ones if kind == 1 else z
        np.testing.assert_allclose(mod.residual(u, zz), simp1.residual(u, ones), rtol=0, atol=1e-13 * np.abs(R0).max())
    if kind == 2:  # saturates above the regularization length
        np.testing.assert_allclose(mod.residual(u, ones * L * 1.01), simp1.residual(u, ones), rtol=0,
                                   atol=1e-13 * np.abs(R0).max())


# ----------------------------------------------------------------- further reference goldens
def test_von_mises_golden(goldens):
    """VonMisesYieldFunction<3> on the two cells of AugLag_VonMises3D
    (PlatoAugLagStressTest.cpp:1063-1101), at the reference's own 1e-4 tolerance (its
    inputs and golds are printed with six decimals)."""
    g = goldens["von_mises3d"]
    got = orc.von_mises(np.array(g["stress"]))
    for v, r in zip(got, g["gold"]):
        assert abs(v - r) <= g["tolerance"] * abs(r), (v, r)
    # a wrong shear factor (1 instead of 3) or a missing 1/2 fails the golden
    s = np.array(g["stress"])
    d = s[:, 0:3] - np.roll(s[:, 0:3], -1, axis=1)
    bad = np.sqrt(0.5 * (d ** 2).sum(1) + 1.0 * (s[:, 3:] ** 2).sum(1))
    assert np.all(np.abs(bad - g["gold"]) > 1e-4 * np.abs(g["gold"]))


def test_volume_golden(box222, goldens):
    """Geometric Volume3D (DerivativeTests.cpp:2151-2302): SIMP exponent 1, minimum 0,
    z = 1 on the 2x2x2 box: value 1, dV/dz (27) and dV/dX (81) at 1e-13."""
    g = goldens["volume3d"]
    model = orc.Model(box222.coords, box222.conn)
    v, gz, gx = model.volume(np.ones(27), model=0, p=1.0, rmin=0.0)
    assert abs(v - g["value_gold"]) <= 1e-13 * g["value_gold"]
    assert_close(gz, g["grad_z_gold"], 1e-13)
    assert_close(gx, g["grad_x_gold"], 1e-13)


def test_stress_pnorm_gradient_x_golden(ref_model, goldens):
    """StressPNorm3D gradient_x (DerivativeTests.cpp:1113-1137): 27 values at p = 12,
    1e-13 as the reference (1e-12 where the gold is printed short)."""
    F, gx = ref_model.pnorm_gradient_x(ref_state(27), np.ones(27), 0, 12.0)
    assert abs(F - goldens["stress_pnorm3d"]["value_gold"]) <= 1e-13 * F
    assert_close(gx[:27], goldens["stress_pnorm3d"]["grad_x_gold"], 1e-12)


def test_pnorm_gradient_x_fd(small):
    """dF/dX of both p-norm measures against central differences of the value in the
    nodal coordinates (interior and boundary nodes)."""
    m, z, u, model, rm, ci = small
    for measure, pn, q in [(0, 8.0, 0.0), (1, 8.0, 0.5)]:
        F, gx = model.pnorm_gradient_x(u, z, measure, pn, q)
        for n in [0, 7, m.num_nodes // 2, m.num_nodes - 1]:
            for k in range(3):
                h = 1e-6
                Xp, Xm = m.coords.copy(), m.coords.copy()
                Xp[n, k] += h
                Xm[n, k] -= h
                mp = orc.Model(Xp, m.conn, E=1.0, nu=0.3, p=3.0, rmin=1e-3)
                mm = orc.Model(Xm, m.conn, E=1.0, nu=0.3, p=3.0, rmin=1e-3)
                fd = (mp.pnorm(u, z, measure, pn, q, False, False)[0]
                      - mm.pnorm(u, z, measure, pn, q, False, False)[0]) / (2 * h)
                assert abs(fd - gx[3 * n + k]) <= 1e-6 * np.abs(gx).max(), (measure, n, k, fd, gx[3 * n + k])


# ----------------------------------------------------------------- penalty weights, closed forms
# Hand-computed weights at the interior density c = 0.5 (and their slopes), from the
# reference's formulas (src/Simp.hpp:130-134, src/Ramp.hpp:94-100, src/Heaviside.hpp:110-125):
#   SIMP p=3, rmin=1e-3:  w = 1e-3 + 0.999 * 0.5^3 = 0.125875;  w' = 0.999 * 3 * 0.25 = 0.74925
#   RAMP p=3, rmin=1e-3:  0.5 / (1 + 3 * 0.5) = 0.2  ->  w = 0.2008;  w' = 0.999 * 4 / 2.5^2 = 0.63936
#   Heaviside p=2, rmin=1e-3, L=1.2: pi * 0.5 / (2 * 1.2) = 37.5 degrees,
#       h = (1 + sin 37.5deg) / 2,  w = 1e-3 + 0.999 h^2,  w' = 0.999 * 2 h * cos(37.5deg) / 2 * pi / 2.4
#   none: w = 1, w' = 0
_S37 = np.sin(np.radians(37.5))
_C37 = np.cos(np.radians(37.5))
_H37 = 0.5 * (1.0 + _S37)
WEIGHTS_AT_HALF = [  # (model, p, rmin, L, w(0.5), w'(0.5))
    (0, 3.0, 1e-3, 1.0, 0.125875, 0.74925),
    (1, 3.0, 1e-3, 1.0, 0.2008, 0.63936),
    (2, 2.0, 1e-3, 1.2, 1e-3 + 0.999 * _H37 ** 2, 0.999 * 2.0 * _H37 * 0.5 * _C37 * np.pi / 2.4),
    (3, 1.0, 0.0, 1.0, 1.0, 0.0),
]


@pytest.mark.parametrize("kind,p,rmin,L,w,dw", WEIGHTS_AT_HALF)
def test_penalty_weight_closed_form_uniform(small, kind, p, rmin, L, w, dw):
    """Uniform z = 0.5 against the unit-weight problem (penalty "none", K1): K(z) = w K1,
    R(u,z) + f = w (R1(u) + f), C(z) = w C1, sum_n dC/dz_n = w' C1, dR/dz summed over
    the control nodes = w' K1 u, the Voigt p-norm scales by w and the volume by w.  A
    dropped (1 - rmin) factor, a wrong exponent or a weight of the mean replaced by the
    mean of the weights fails here."""
    m, z, u, _, rm, ci = small
    fv, fe = faces_on_plane(m, 0, 2.0)
    t = (0.0, 0.0, -1.0)
    mod = orc.Model(m.coords, m.conn, E=1.0, nu=0.3, p=p, rmin=rmin, face_verts=fv, traction=t,
                    penalty_model=kind, length=L)
    unit = orc.Model(m.coords, m.conn, E=1.0, nu=0.3, p=1.0, rmin=0.0, face_verts=fv, traction=t,
                     penalty_model=3)
    half, one = np.full(m.num_nodes, 0.5), np.ones(m.num_nodes)
    K1 = unit.jacobian(u, one, rm, ci)
    Kc = mod.jacobian(u, half, rm, ci)
    np.testing.assert_allclose(Kc, w * K1, rtol=0, atol=1e-13 * np.abs(K1).max())
    f = -mod.residual(np.zeros_like(u), one)
    np.testing.assert_allclose(mod.residual(u, half) + f, w * (unit.residual(u, one) + f), rtol=0,
                               atol=1e-13 * np.abs(K1).max() * np.abs(u).max())
    C1 = unit.energy(u, one, False, False)[0]
    Cc, _, gz = mod.energy(u, half, False, True)
    assert abs(Cc - w * C1) <= 1e-13 * C1
    assert abs(gz.sum() - dw * C1) <= 1e-12 * max(abs(dw) * C1, 1e-300) + 1e-15 * C1
    G = mod.gradient_z(u, half, rm, ci)  # rows = control node; entries dR_(col,i)/dz_row
    colsum = np.zeros(3 * m.num_nodes)
    for r in range(m.num_nodes):
        for s in range(rm[r], rm[r + 1]):
            colsum[3 * ci[s]:3 * ci[s] + 3] += G[3 * s:3 * s + 3]
    K1u = orc.to_dense(rm, ci, K1) @ u
    np.testing.assert_allclose(colsum, dw * K1u, rtol=0, atol=1e-12 * np.abs(K1u).max())
    F1 = unit.pnorm(u, one, 0, 8.0, 0.0, False, False)[0]
    Fc = mod.pnorm(u, half, 0, 8.0, 0.0, False, False)[0]
    assert abs(Fc - w * F1) <= 1e-13 * F1
    box = 2.0
    assert abs(mod.volume(half, kind, p, rmin, L)[0] - w * box) <= 1e-13 * box


@pytest.mark.parametrize("kind,p,rmin,L", [(0, 3.0, 1e-3, 1.0), (1, 3.0, 1e-3, 1.0), (2, 2.0, 1e-3, 1.2)])
def test_penalty_weight_of_mean_density_one_element(kind, p, rmin, L):
    """One tet with distinct nodal densities (mean 0.525): its stiffness is
    w(mean z) K(z = 1) (ApplyWeighting averages the four nodal controls before the
    penalty, src/ApplyWeighting.hpp:49-65)."""
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    conn = np.array([[0, 1, 2, 3]], np.int32)
    rm, ci = orc.graph(conn, 4)
    mod = orc.Model(X, conn, E=1.0, nu=0.3, p=p, rmin=rmin, penalty_model=kind, length=L)
    zr = np.array([0.2, 0.4, 0.6, 0.9])
    rho = 0.525
    if kind == 0:
        w = 1e-3 + 0.999 * rho ** 3
    elif kind == 1:
        w = 1e-3 + 0.999 * rho / (1.0 + 3.0 * (1.0 - rho))
    else:
        w = 1e-3 + 0.999 * (0.5 * (1.0 + np.sin(np.pi * rho / 2.4))) ** 2
    K1 = orc.Model(X, conn, E=1.0, nu=0.3, p=1.0, rmin=0.0, penalty_model=3).jacobian(None, np.ones(4), rm, ci)
    np.testing.assert_allclose(mod.jacobian(None, zr, rm, ci), w * K1, rtol=0, atol=1e-14 * np.abs(K1).max())


def test_von_mises_pnorm_density_scaling(small):
    """Relaxed von Mises p-norm: uniform z = c scales F by c^q exactly (0.5^0.5), and
    the measure of one element equals rho^q times the von Mises stress of its
    unpenalized stress (checked through the p = 1 norm of a one-element mesh)."""
    m, z, u, model, rm, ci = small
    F1 = model.pnorm(u, np.ones(m.num_nodes), 1, 8.0, 0.5, False, False)[0]
    Fc = model.pnorm(u, np.full(m.num_nodes, 0.5), 1, 8.0, 0.5, False, False)[0]
    assert abs(Fc - np.sqrt(0.5) * F1) <= 1e-13 * F1
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    conn = np.array([[0, 1, 2, 3]], np.int32)
    one = orc.Model(X, conn, E=1.0, nu=0.3, p=3.0, rmin=1e-3)
    ue = np.random.default_rng(3).normal(size=12) * 1e-2
    zr = np.array([0.2, 0.4, 0.6, 0.9])
    _, vol, eps, sig, _ = one.cell_worksets(ue, 1)
    vm = orc.von_mises(sig)[0]
    F = one.pnorm(ue, zr, 1, 1.0, 0.5, False, False)[0]
    assert abs(F - vol[0] * 0.525 ** 0.5 * vm) <= 1e-14 * F
