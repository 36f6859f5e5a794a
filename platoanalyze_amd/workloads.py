"""Synthetic SIMP cantilever workloads (the measured configurations).

Box [0,2]x[0,1]x[0,1] split into nx*ny*nz hexes x 6 Tet4; isotropic E=1, nu=0.3;
MSIMP p=3, rho_min=1e-3; node set x=0 fixed in all three dofs; uniform traction
(0,0,-1) on the side set x=2.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .mesh import TetMesh, build_box, faces_on_plane, nodes_on_plane, smoothed_density


@dataclass(frozen=True)
class Config:
    name: str
    nx: int
    ny: int
    nz: int
    density: str  # "uniform" (0.5) or "smoothed" (seeded, filtered, sigmoid)
    lx: float = 2.0
    ly: float = 1.0
    lz: float = 1.0
    E: float = 1.0
    nu: float = 0.3
    p: float = 3.0
    rho_min: float = 1e-3
    seed: int = 20260101


CONFIGS = {
    "cantilever_20x10x10": Config("cantilever_20x10x10", 20, 10, 10, "uniform"),
    "cantilever_128x64x64": Config("cantilever_128x64x64", 128, 64, 64, "smoothed"),
    "cantilever_64x32x32": Config("cantilever_64x32x32", 64, 32, 32, "smoothed"),
    "cantilever_256x128x128": Config("cantilever_256x128x128", 256, 128, 128, "smoothed"),
}


@dataclass
class Workload:
    cfg: Config
    mesh: TetMesh
    fixed_nodes: np.ndarray
    face_verts: np.ndarray
    face_elems: np.ndarray
    traction: tuple
    z: np.ndarray

    @property
    def bc_dofs(self) -> np.ndarray:
        return np.sort(np.concatenate([3 * self.fixed_nodes + d for d in range(3)])).astype(np.int32)


def make(cfg: Config | str, tiny: tuple | None = None) -> Workload:
    if isinstance(cfg, str):
        cfg = CONFIGS[cfg]
    if tiny is not None:  # same physics on a smaller grid
        cfg = Config(f"{cfg.name}@{tiny[0]}x{tiny[1]}x{tiny[2]}", *tiny, cfg.density, cfg.lx, cfg.ly, cfg.lz,
                     cfg.E, cfg.nu, cfg.p, cfg.rho_min, cfg.seed)
    mesh = build_box(cfg.nx, cfg.ny, cfg.nz, cfg.lx, cfg.ly, cfg.lz)
    tol = 1e-9 * cfg.lx
    fixed = nodes_on_plane(mesh, 0, 0.0, tol)
    fv, fe = faces_on_plane(mesh, 0, cfg.lx, tol)
    if cfg.density == "uniform":
        z = np.full(mesh.num_nodes, 0.5)
    else:
        z = smoothed_density(cfg.nx, cfg.ny, cfg.nz, mesh, seed=cfg.seed, rho_min=cfg.rho_min)
    return Workload(cfg, mesh, fixed, fv, fe, (0.0, 0.0, -1.0), z)


def model(w: Workload, device: str = "cuda", graph=None):
    """ElasticModel (device handle) of a workload."""
    from .elliptic import ElasticModel, EssentialBC, Traction

    bcs = [EssentialBC(w.fixed_nodes, d, 0.0) for d in range(3)]
    tr = [Traction(w.face_verts, w.face_elems, w.traction)]
    return ElasticModel(w.mesh, w.cfg.E, w.cfg.nu, w.cfg.p, w.cfg.rho_min, bcs, tr, graph=graph, device=device,
                        box=(w.cfg.nx, w.cfg.ny, w.cfg.nz))

