"""tests/golden/oc_design_64x32x32.npz: a late SIMP optimality-criteria design on the
64x32x32 x 6 Tet4 cantilever whose block-Jacobi spectrum has its top on a few solid/void
interfaces.

    python tests/golden/make_oc_design.py generate   # GPU: writes the fixture
    python tests/golden/make_oc_design.py spectrum   # CPU: lambda_max(D^-1 K) of the fixture

`generate` runs the fourth trajectory of the warm-start training data (numpy
default_rng(11), three (vf, d) pairs skipped: volume fraction vf, unit traction d on the
x = L face) for 76 OC steps, each solved by mixed-precision multigrid PCG as that data
generation does, and stores the design after step 76.  The committed fixture came from the
build whose multigrid smoother took its interval from a 12-step power iteration; solves
differ in rounding from build to build, so a rerun gives a nearby design of the same kind.

`spectrum` assembles K(z) with the oracle, restricts it to the free dofs, and prints the
top three eigenvalues of D^-1 K (D its 3x3 block diagonal; scipy eigsh), the 12-step
power-iteration estimate and the 16-step PCG-Lanczos estimate from the same start vector.
For the committed fixture: top 2.4029, 2.4504, 2.6291; power 2.108; Lanczos 2.6254.
"""
import dataclasses
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
FIXTURE = Path(__file__).with_name("oc_design_64x32x32.npz")


def generate(steps: int = 76) -> None:
    import torch

    from platoanalyze_amd import workloads as wl
    from platoanalyze_amd.elliptic import Problem
    from platoanalyze_amd.warmstart import nodal_volumes, oc_update

    base = wl.make("cantilever_64x32x32")
    rng = np.random.default_rng(11)
    for _ in range(4):
        vf = float(rng.uniform(0.3, 0.5))
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
    w = dataclasses.replace(base, traction=tuple(float(v) for v in d))
    m = wl.model(w)
    pb = Problem(m, tolerance=1e-8, check_every=4, precision="mixed", preconditioner="multigrid")
    wn = nodal_volumes(base)
    z = np.full(base.mesh.num_nodes, vf)
    for _ in range(steps):
        zt = torch.as_tensor(z, device="cuda")
        u = pb.solution(zt)
        g, _ = pb.criterionGradient(zt, u)
        z = oc_update(z, g.cpu().numpy(), wn, vf, base.cfg.rho_min)
    np.savez_compressed(FIXTURE, z=z, traction=np.asarray(d), vf=vf, steps=steps)
    print(f"wrote {FIXTURE}: vf {vf:.4f}, traction {np.round(d, 4)}")


def spectrum() -> None:
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    from scipy.linalg import eigh_tridiagonal

    from oracle import oracle as orc
    from platoanalyze_amd import workloads as wl

    w = wl.make("cantilever_64x32x32")
    c, N = w.cfg, w.mesh.num_nodes
    z = np.load(FIXTURE)["z"]
    o = orc.Model(w.mesh.coords, w.mesh.conn, c.E, c.nu, c.p, c.rho_min)
    rm, ci = orc.graph(w.mesh.conn, N)
    v = np.asarray(o.jacobian(np.zeros(3 * N), z, rm, ci)).reshape(-1, 3, 3)
    fixed = np.abs(w.mesh.coords[:, 0]) < 1e-12
    free_nodes = np.where(~fixed)[0]
    idx = (3 * free_nodes[:, None] + np.arange(3)).ravel()
    K = sp.bsr_matrix((v, ci, rm), shape=(3 * N, 3 * N)).tocsr()[idx][:, idx].tocsr()
    diag = np.array([v[k] for r in range(N) for k in range(rm[r], rm[r + 1]) if ci[k] == r])[free_nodes]
    Dinv = sp.block_diag(list(np.linalg.inv(diag)), format="csr")
    Li = sp.block_diag(list(np.linalg.inv(np.linalg.cholesky(diag))), format="csr")
    top = np.sort(sla.eigsh(Li @ K @ Li.T, k=3, which="LA", return_eigenvectors=False, tol=1e-6))
    n = K.shape[0]
    start = ((np.arange(n, dtype=np.uint64) * np.uint64(2654435761) >> np.uint64(7)) % np.uint64(2001)) / 1000.0 - 1.0
    x = start.copy()
    A = Li @ K @ Li.T
    for _ in range(12):
        x = A @ (x / np.linalg.norm(x))
    r = start.copy()
    zz = Dinv @ r
    p, rz, al, be, ta, tb = zz.copy(), r @ zz, [], [], [], []
    for j in range(16):
        q = K @ p
        a = rz / (p @ q)
        r = r - a * q
        zz = Dinv @ r
        rz1 = r @ zz
        ta.append(1 / a + (be[-1] / al[-1] if j else 0.0))
        if j:
            tb.append(np.sqrt(be[-1]) / al[-1])
        al.append(a)
        be.append(rz1 / rz)
        p, rz = zz + be[-1] * p, rz1
    lz = eigh_tridiagonal(np.array(ta), np.array(tb), eigvals_only=True).max()
    print(f"top {np.round(top, 4)}; power(12) {np.linalg.norm(x):.4f}; Lanczos(16) {lz:.4f}")


if __name__ == "__main__":
    {"generate": generate, "spectrum": spectrum}[sys.argv[1] if len(sys.argv) > 1 else "spectrum"]()
