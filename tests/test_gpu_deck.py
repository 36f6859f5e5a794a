"""Deck-built problems and the MPMD operation surface on the GPU.

* A deck with the reference's DerivativeTests parameters (E = 1e6, nu = 0.3, SIMP p = 1,
  minimum 0) reproduces its residual and internal-energy goldens on the 2x2x2 box.
* The block product y += alpha A x (pa_matvec_plus_vector) against a dense product of the
  same blocks, for 3x3 and 1x3 blocks.
* Criterion gradients: the non-self-adjoint path (adjoint solve, (dR/dz)^T lambda and
  (dR/dX)^T lambda through the block product) against the self-adjoint one in z, against
  central differences of the oracle's compliance over nodal coordinates in X, and the
  generic adjoint against the fused p-norm adjoint.
* An app file and an Exodus mesh drive ComputeSolution / ComputeCriterion* operations,
  with values and gradients exported by ArgumentName.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from platoanalyze_amd import _capi
from platoanalyze_amd import workloads as wl
from platoanalyze_amd.app import MPMDApp, parse_operations
from platoanalyze_amd.deck import ElasticProblem, parse_parameter_list
from platoanalyze_amd.elliptic import PhysicsScalarFunction, Problem, VectorFunction, _dptr
from platoanalyze_amd.exodus import ExodusMesh, read_exodus, sides_of_faces, write_exodus

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def rel(got, ref):
    return np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300)


def ref_state(nn):
    return 1e-4 * np.arange(1, 3 * nn + 1, dtype=np.float64)


GOLDEN_DECK = """
<ParameterList name='Plato Problem'>
  <ParameterList name='Spatial Model'>
    <ParameterList name='Domains'>
      <ParameterList name='Design Volume'>
        <Parameter name='Element Block' type='string' value='body'/>
        <Parameter name='Material Model' type='string' value='Unobtainium'/>
      </ParameterList>
    </ParameterList>
  </ParameterList>
  <Parameter name='PDE Constraint' type='string' value='Elliptic'/>
  <Parameter name='Self-Adjoint' type='bool' value='true'/>
  <ParameterList name='Elliptic'>
    <ParameterList name='Penalty Function'>
      <Parameter name='Type' type='string' value='SIMP'/>
      <Parameter name='Exponent' type='double' value='1.0'/>
      <Parameter name='Minimum Value' type='double' value='0.0'/>
    </ParameterList>
  </ParameterList>
  <ParameterList name='Material Models'>
    <ParameterList name='Unobtainium'>
      <ParameterList name='Isotropic Linear Elastic'>
        <Parameter name='Poissons Ratio' type='double' value='0.3'/>
        <Parameter name='Youngs Modulus' type='double' value='1.0e6'/>
      </ParameterList>
    </ParameterList>
  </ParameterList>
  <ParameterList name='Criteria'>
    <ParameterList name='Internal Elastic Energy'>
      <Parameter name='Type' type='string' value='Scalar Function'/>
      <Parameter name='Scalar Function Type' type='string' value='Internal Elastic Energy'/>
      <ParameterList name='Penalty Function'>
        <Parameter name='Type' type='string' value='SIMP'/>
        <Parameter name='Exponent' type='double' value='1.0'/>
        <Parameter name='Minimum Value' type='double' value='0.0'/>
      </ParameterList>
    </ParameterList>
  </ParameterList>
</ParameterList>
"""


def test_gpu_deck_reference_goldens_2x2x2(box222, goldens):
    print("case: deck goldens 2x2x2", flush=True)
    ex = ExodusMesh(box222, {}, {}, {"body": (0, box222.num_elems)})
    pb = ElasticProblem(parse_parameter_list(GOLDEN_DECK), ex)
    assert pb.self_adjoint and pb.E == 1e6 and pb.penalty == ("SIMP", 1.0, 0.0, 1.0)
    assert not pb.criteria["Internal Elastic Energy"].own_handle
    u, z = dev(ref_state(27)), dev(np.ones(27))
    g = goldens["elastostatic_residual3d"]
    r = host(VectorFunction(pb.model).value(u, z))
    assert rel(r, np.asarray(g["residual_gold"])) <= 1e-13
    e = goldens["internal_elastic_energy3d"]
    v = pb.criterionValue(z, u, "Internal Elastic Energy")
    assert abs(v - e["value_gold"]) <= 1e-13 * e["value_gold"]
    # self-adjoint criterionGradient = -dC/dz|_u
    assert rel(host(pb.criterionGradient(z, u, "Internal Elastic Energy")), -np.asarray(e["grad_z_gold"])) <= 1e-13


@pytest.mark.parametrize("block_rows", [3, 1])
def test_gpu_matvec_plus_vector(block_rows):
    print(f"case: matvec_plus_vector {block_rows}x3", flush=True)
    w = wl.make("cantilever_20x10x10", tiny=(5, 3, 2))
    m = wl.model(w)
    rm, ci = host(m.row_map), host(m.column_indices)
    rng = np.random.default_rng(block_rows)
    vals = rng.normal(size=3 * block_rows * m.nnz_blocks)
    x = rng.normal(size=3 * m.num_nodes)
    y0 = rng.normal(size=block_rows * m.num_nodes)
    A = np.zeros((block_rows * m.num_nodes, 3 * m.num_nodes))
    vb = vals.reshape(-1, block_rows, 3)
    for r in range(m.num_nodes):
        for s in range(rm[r], rm[r + 1]):
            A[block_rows * r:block_rows * (r + 1), 3 * ci[s]:3 * ci[s] + 3] += vb[s]
    y, dvals, dx = dev(y0), dev(vals), dev(x)  # held: a freed temporary's memory is reused
    m._stream()
    _capi.call("pa_matvec_plus_vector", m.handle, block_rows, -0.5, _dptr(dvals), _dptr(dx), _dptr(y))
    ref = y0 - 0.5 * A @ x
    np.testing.assert_allclose(host(y), ref, rtol=0, atol=1e-13 * np.abs(ref).max())
    with pytest.raises(_capi.PaError):
        _capi.call("pa_matvec_plus_vector", m.handle, 2, 1.0, _dptr(dvals), _dptr(dx), _dptr(y))


def _cantilever_deck(self_adjoint: bool, extra_criteria: str = "", solver: str = "", tol: float = 1e-12) -> str:
    return f"""
<ParameterList name='Problem'>
  <Parameter name='Spatial Dimension' type='int' value='3'/>
  <ParameterList name='Plato Problem'>
    <Parameter name='Physics' type='string' value='Mechanical'/>
    <Parameter name='PDE Constraint' type='string' value='Elliptic'/>
    <Parameter name='Self-Adjoint' type='bool' value='{str(self_adjoint).lower()}'/>
    <ParameterList name='Elliptic'>
      <ParameterList name='Penalty Function'>
        <Parameter name='Type' type='string' value='SIMP'/>
        <Parameter name='Exponent' type='double' value='3.0'/>
        <Parameter name='Minimum Value' type='double' value='1.0e-3'/>
      </ParameterList>
    </ParameterList>
    <ParameterList name='Material Models'>
      <ParameterList name='steel'>
        <ParameterList name='Isotropic Linear Elastic'>
          <Parameter name='Youngs Modulus' type='double' value='1.0'/>
          <Parameter name='Poissons Ratio' type='double' value='0.3'/>
        </ParameterList>
      </ParameterList>
    </ParameterList>
    <ParameterList name='Essential Boundary Conditions'>
      <ParameterList name='X Fixed'>
        <Parameter name='Type' type='string' value='Zero Value'/>
        <Parameter name='Index' type='int' value='0'/>
        <Parameter name='Sides' type='string' value='x-'/>
      </ParameterList>
      <ParameterList name='Y Fixed'>
        <Parameter name='Type' type='string' value='Zero Value'/>
        <Parameter name='Index' type='int' value='1'/>
        <Parameter name='Sides' type='string' value='x-'/>
      </ParameterList>
      <ParameterList name='Z Fixed'>
        <Parameter name='Type' type='string' value='Fixed Value'/>
        <Parameter name='Index' type='int' value='2'/>
        <Parameter name='Value' type='double' value='0.0'/>
        <Parameter name='Sides' type='string' value='x-'/>
      </ParameterList>
    </ParameterList>
    <ParameterList name='Natural Boundary Conditions'>
      <ParameterList name='Tip Load'>
        <Parameter name='Type' type='string' value='Uniform'/>
        <Parameter name='Index' type='int' value='2'/>
        <Parameter name='Value' type='double' value='-1.0'/>
        <Parameter name='Sides' type='string' value='x+'/>
      </ParameterList>
    </ParameterList>
    <ParameterList name='Criteria'>
      <ParameterList name='My Compliance'>
        <Parameter name='Type' type='string' value='Scalar Function'/>
        <Parameter name='Scalar Function Type' type='string' value='Internal Elastic Energy'/>
        <ParameterList name='Penalty Function'>
          <Parameter name='Type' type='string' value='SIMP'/>
          <Parameter name='Exponent' type='double' value='3.0'/>
          <Parameter name='Minimum Value' type='double' value='1.0e-3'/>
        </ParameterList>
      </ParameterList>
      <ParameterList name='My Volume'>
        <Parameter name='Type' type='string' value='Scalar Function'/>
        <Parameter name='Scalar Function Type' type='string' value='Volume'/>
        <Parameter name='Linear' type='bool' value='true'/>
      </ParameterList>
      {extra_criteria}
    </ParameterList>
    <ParameterList name='Linear Solver'>
      <Parameter name='Solver Stack' type='string' value='AmgX'/>
      <Parameter name='Tolerance' type='double' value='{tol}'/>
      {solver}
    </ParameterList>
  </ParameterList>
</ParameterList>
"""


def _cantilever_mesh(dims=(4, 2, 2)):
    w = wl.make("cantilever_20x10x10", tiny=dims)
    el, sd = sides_of_faces(w.mesh, w.face_verts, w.face_elems)
    ex = ExodusMesh(w.mesh, {"x-": w.fixed_nodes}, {"x+": (w.face_verts, w.face_elems)},
                    {"body": (0, w.mesh.num_elems)})
    z = np.clip(w.z + 0.3 * np.random.default_rng(5).uniform(-1, 1, w.mesh.num_nodes), 0.05, 1.0)
    return w, ex, z, (el, sd)


def test_gpu_deck_matches_direct_api():
    print("case: deck cantilever vs direct API", flush=True)
    w, ex, z, _ = _cantilever_mesh()
    pb = ElasticProblem(parse_parameter_list(_cantilever_deck(True)), ex)
    ref = wl.model(w)
    dz = dev(z)
    u = pb.solution(dz)
    u_ref = Problem(ref, tolerance=1e-12).solution(dz)
    assert rel(host(u), host(u_ref)) <= 1e-10
    v = pb.criterionValue(dz, u, "My Compliance")
    assert abs(v - PhysicsScalarFunction(ref).value(u_ref, dz)) <= 1e-10 * abs(v)
    assert abs(pb.criterionValue(dz, u, "My Volume") - 2.0) <= 1e-13  # NoPenalty volume of the 2x1x1 box
    assert pb.criterionIsLinear("My Volume") and not pb.criterionIsLinear("My Compliance")
    with pytest.raises(KeyError):
        pb.criterionValue(dz, u, "nope")


def _box_active(m) -> int:
    import ctypes as C
    v = C.c_int(-1)
    _capi.call("pa_box_assembly", m.handle, C.byref(v))
    return v.value


def test_gpu_deck_declares_box_for_every_solver():
    """A box mesh is declared whatever the deck's solver, so the lattice assembly applies
    under block Jacobi too; a tensor lattice with unequal spacings is no box: a Jacobi deck
    still builds (general assembly) and matches the direct API, a multigrid deck is refused."""
    print("case: deck box declaration", flush=True)
    from platoanalyze_amd.elliptic import ElasticModel, EssentialBC, Traction
    from platoanalyze_amd.mesh import TetMesh
    w, ex, z, _ = _cantilever_mesh()
    dz = dev(z)
    pb = ElasticProblem(parse_parameter_list(_cantilever_deck(True)), ex)
    assert pb.box == (4, 2, 2) and _box_active(pb.model) == 1
    X = w.mesh.coords.copy()
    X[:, 0] = X[:, 0] ** 1.3 / 2.0 ** 0.3  # graded along x, still spanning [0, 2]
    graded = TetMesh(X, w.mesh.conn)
    ex2 = ExodusMesh(graded, ex.node_sets, ex.side_sets, {"body": (0, graded.num_elems)})
    pb2 = ElasticProblem(parse_parameter_list(_cantilever_deck(True)), ex2)
    assert pb2.box is None and _box_active(pb2.model) == 0
    bcs = [EssentialBC(w.fixed_nodes, d, 0.0) for d in range(3)]
    ref = ElasticModel(graded, 1.0, 0.3, 3.0, 1e-3, bcs, [Traction(w.face_verts, w.face_elems, (0.0, 0.0, -1.0))])
    assert rel(host(pb2.solution(dz)), host(Problem(ref, tolerance=1e-12).solution(dz))) <= 1e-10
    mg = _cantilever_deck(True, solver="<Parameter name='Preconditioner' type='string' value='Multigrid'/>")
    with pytest.raises(_capi.PaError):
        ElasticProblem(parse_parameter_list(mg), ex2)


def _oracle_compliance(coords, w, z, rm, ci):
    o = orc.Model(coords, w.mesh.conn, w.cfg.E, w.cfg.nu, w.cfg.p, w.cfg.rho_min, w.face_verts, w.traction)
    f = -o.residual(np.zeros(3 * w.mesh.num_nodes), z)
    Kc, bc = orc.apply_constraints(rm, ci, o.jacobian(None, z, rm, ci), f, w.bc_dofs, np.zeros(len(w.bc_dofs)))
    u = np.linalg.solve(orc.to_dense(rm, ci, Kc), bc)
    return o.energy(u, z, False, False)[0]


def _oracle_pnorm(coords, w, z, rm, ci, pn):
    """Voigt stress p-norm (SIMP p = 3, minimum 1e-3) of the oracle's dense solve."""
    o = orc.Model(coords, w.mesh.conn, w.cfg.E, w.cfg.nu, w.cfg.p, w.cfg.rho_min, w.face_verts, w.traction)
    f = -o.residual(np.zeros(3 * w.mesh.num_nodes), z)
    Kc, bc = orc.apply_constraints(rm, ci, o.jacobian(None, z, rm, ci), f, w.bc_dofs, np.zeros(len(w.bc_dofs)))
    u = np.linalg.solve(orc.to_dense(rm, ci, Kc), bc)
    return o.pnorm(u, z, 0, pn, 0.0, False, False)[0]


def test_gpu_criterion_gradients_non_self_adjoint():
    """Non-self-adjoint dC/dz equals the self-adjoint -dC/dz|_u (lambda = -u under zero
    Dirichlet values); non-self-adjoint dC/dX equals central differences (h = 1e-6) of the
    oracle's compliance of dense solves at displaced coordinates, including the traction
    face term that the reference's self-adjoint branch (-dC/dX|_u) leaves out."""
    print("case: deck criterion gradients", flush=True)
    w, ex, z, _ = _cantilever_mesh((3, 2, 2))
    dz = dev(z)
    sa = ElasticProblem(parse_parameter_list(_cantilever_deck(True)), ex)
    na = ElasticProblem(parse_parameter_list(_cantilever_deck(False)), ex)
    u = na.solution(dz)
    gz_sa = host(sa.criterionGradient(dz, u, "My Compliance"))
    gz_na = host(na.criterionGradient(dz, u, "My Compliance"))
    assert rel(gz_na, gz_sa) <= 1e-9
    gx = host(na.criterionGradientX(dz, u, "My Compliance"))
    gx_sa = host(sa.criterionGradientX(dz, u, "My Compliance"))
    o = orc.Model(w.mesh.coords, w.mesh.conn, w.cfg.E, w.cfg.nu, w.cfg.p, w.cfg.rho_min, w.face_verts, w.traction)
    assert rel(gx_sa, -o.energy_gradient_x(host(u), z)) <= 1e-12
    rm, ci = host(na.model.row_map), host(na.model.column_indices)
    h = 1e-6
    for n in range(w.mesh.num_nodes):
        for k in range(3):
            vals = []
            for s in (1.0, -1.0):
                X = w.mesh.coords.copy()
                X[n, k] += s * h
                vals.append(_oracle_compliance(X, w, z, rm, ci))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - gx[3 * n + k]) <= 1e-6 * np.abs(gx).max(), (n, k, fd, gx[3 * n + k])
    # the volume criterion's dV/dX through the problem
    vol = host(na.criterionGradientX(dz, u, "My Volume"))
    assert abs(vol.sum()) <= 1e-12  # a rigid translation keeps the volume


PNORM_CRITERIA = """
      <ParameterList name='Stress'>
        <Parameter name='Type' type='string' value='Scalar Function'/>
        <Parameter name='Scalar Function Type' type='string' value='Stress P-Norm'/>
        <Parameter name='Exponent' type='double' value='6.0'/>
        <ParameterList name='Penalty Function'>
          <Parameter name='Type' type='string' value='SIMP'/>
          <Parameter name='Exponent' type='double' value='3.0'/>
          <Parameter name='Minimum Value' type='double' value='1.0e-3'/>
        </ParameterList>
      </ParameterList>
      <ParameterList name='Stress RAMP'>
        <Parameter name='Type' type='string' value='Scalar Function'/>
        <Parameter name='Scalar Function Type' type='string' value='Stress P-Norm'/>
        <Parameter name='Exponent' type='double' value='6.0'/>
        <ParameterList name='Penalty Function'>
          <Parameter name='Type' type='string' value='RAMP'/>
          <Parameter name='Exponent' type='double' value='2.0'/>
        </ParameterList>
      </ParameterList>
"""


def test_gpu_pnorm_generic_adjoint_matches_fused():
    """The p-norm criterion on the residual's penalty uses the fused adjoint
    (pa_stress_pnorm_gradient); forced onto the generic path (assembled K, adjoint
    constraints, PCG, (dR/dz)^T lambda by the block product) it gives the same gradient.
    A criterion with its own penalty (RAMP) is evaluated on a second handle, checked
    against central differences of its value over two controls.  The configuration
    gradient of the p-norm (adjoint with (dR/dX)^T lambda) against central differences of
    the oracle's p-norm over nodal coordinates."""
    print("case: deck p-norm adjoints", flush=True)
    w, ex, z, _ = _cantilever_mesh((4, 2, 2))
    dz = dev(z)
    pb = ElasticProblem(parse_parameter_list(_cantilever_deck(False, PNORM_CRITERIA)), ex)
    u = pb.solution(dz)
    c = pb.criteria["Stress"]
    assert not c.own_handle and pb.criteria["Stress RAMP"].own_handle
    g_fused = host(pb.criterionGradient(dz, u, "Stress"))
    c.own_handle = True
    g_generic = host(pb.criterionGradient(dz, u, "Stress"))
    assert rel(g_generic, g_fused) <= 1e-8
    g = host(pb.criterionGradient(dz, u, "Stress RAMP"))
    h = 1e-6
    for n in (int(np.argmax(np.abs(g))), w.mesh.num_nodes // 2):
        vals = []
        for s in (1.0, -1.0):
            zz = z.copy()
            zz[n] += s * h
            d = dev(zz)
            vals.append(pb.criterionValue(d, pb.solution(d), "Stress RAMP"))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - g[n]) <= 1e-6 * np.abs(g).max() + 1e-5 * abs(fd), (n, fd, g[n])
    # dF/dX of the p-norm through the adjoint: central differences (h = 1e-6) of the
    # oracle's p-norm of dense solves at displaced coordinates
    gx = host(pb.criterionGradientX(dz, u, "Stress"))
    rm, ci = host(pb.model.row_map), host(pb.model.column_indices)
    for n in (int(np.argmax(np.abs(gx)) // 3), w.mesh.num_nodes // 2, int(w.face_verts[0, 0])):
        for k in range(3):
            vals = []
            for s in (1.0, -1.0):
                X = w.mesh.coords.copy()
                X[n, k] += s * h
                vals.append(_oracle_pnorm(X, w, z, rm, ci, 6.0))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - gx[3 * n + k]) <= 1e-6 * np.abs(gx).max(), (n, k, fd, gx[3 * n + k])


APP = """
<root>
  <Operation><Function>ComputeSolution</Function><Name>Solve</Name></Operation>
  <Operation>
    <Function>ComputeCriterion</Function><Name>Objective</Name><Criterion>My Compliance</Criterion>
    <Output><Argument>Value</Argument><ArgumentName>Objective Value</ArgumentName></Output>
    <Output><Argument>Gradient</Argument><ArgumentName>Objective Gradient</ArgumentName></Output>
  </Operation>
  <Operation>
    <Function>ComputeCriterionValue</Function><Name>Constraint</Name><Criterion>My Volume</Criterion>
    <Target>0.5</Target>
    <Output><Argument>Value</Argument><ArgumentName>Constraint Value</ArgumentName></Output>
  </Operation>
  <Operation>
    <Function>ComputeCriterionGradient</Function><Name>Constraint Gradient</Name><Criterion>My Volume</Criterion>
    <Output><Argument>Gradient</Argument><ArgumentName>Constraint Gradient</ArgumentName></Output>
  </Operation>
  <Operation>
    <Function>ComputeCriterionGradientX</Function><Name>Objective X</Name><Criterion>My Compliance</Criterion>
    <Output><Argument>Gradient</Argument><ArgumentName>Objective Gradient X</ArgumentName></Output>
  </Operation>
  <Operation><Function>UpdateProblem</Function><Name>Update</Name></Operation>
  <Operation><Function>Reinitialize</Function><Name>Reinit</Name></Operation>
</root>
"""


def test_gpu_mpmd_app(tmp_path):
    print("case: MPMD app", flush=True)
    w, ex, z, sides = _cantilever_mesh((8, 4, 4))
    mesh_path = tmp_path / "cantilever.exo"
    write_exodus(mesh_path, w.mesh, {"x-": w.fixed_nodes}, {"x+": sides})
    deck = _cantilever_deck(True, solver="<Parameter name='Preconditioner' type='string' value='Multigrid'/>"
                                         "<Parameter name='Precision' type='string' value='Mixed'/>", tol=1e-11)
    deck = deck.replace("<Parameter name='Spatial Dimension' type='int' value='3'/>",
                        f"<Parameter name='Spatial Dimension' type='int' value='3'/>"
                        f"<Parameter name='Input Mesh' type='string' value='{mesh_path}'/>")
    (tmp_path / "deck.xml").write_text(deck)
    (tmp_path / "app.xml").write_text(APP)
    app = MPMDApp(tmp_path / "deck.xml", tmp_path / "app.xml")
    assert app.problem.box == (8, 4, 4)
    assert float(host(app.control).min()) == 1.0  # the control starts at 1
    app.importData("Topology", z)
    app.compute("Solve")
    ref = wl.model(w)
    dz = dev(z)
    u_ref = Problem(ref, tolerance=1e-12).solution(dz)
    assert rel(host(app.exportData("Solution")), host(u_ref)) <= 1e-9
    app.compute("Objective")
    C = PhysicsScalarFunction(ref).value(u_ref, dz)
    assert abs(app.exportData("Objective Value") - C) <= 1e-9 * C
    g_ref, _ = Problem(ref, tolerance=1e-12).criterionGradient(dz, u_ref)
    assert rel(host(app.exportData("Objective Gradient")), host(g_ref)) <= 1e-8
    app.compute("Constraint")
    assert abs(app.exportData("Constraint Value") - (2.0 - 0.5)) <= 1e-13
    app.compute("Constraint Gradient")
    assert abs(float(host(app.exportData("Constraint Gradient")).sum())) <= 1e-13  # NoPenalty: dV/dz = 0
    app.compute("Objective X")
    assert app.exportData("Objective Gradient X").shape == (3 * w.mesh.num_nodes,)
    app.compute("Update")
    app.compute("Reinit")
    assert rel(host(app.control), z) == 0.0
    with pytest.raises(KeyError):
        app.compute("missing")
    with pytest.raises(KeyError):
        app.exportData("missing")
    with pytest.raises(KeyError):  # an operation naming an undefined criterion fails at set-up
        MPMDApp(tmp_path / "deck.xml", parse_operations(APP.replace("My Volume", "Other")), mesh=read_exodus(mesh_path))
