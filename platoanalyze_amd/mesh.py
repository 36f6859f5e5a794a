"""Host-side mesh helpers over the C ABI (no GPU needed).

``build_box`` reproduces the vertex numbering of ``Omega_h::build_box`` (the mesh the
reference tests use, unit_tests/PlatoTestHelpers.hpp:206-250).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _capi


@dataclass
class TetMesh:
    coords: np.ndarray  # (N,3) float64
    conn: np.ndarray    # (E,4) int32

    @property
    def num_nodes(self) -> int:
        return int(self.coords.shape[0])

    @property
    def num_elems(self) -> int:
        return int(self.conn.shape[0])


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def build_box(nx: int, ny: int, nz: int, lx: float = 1.0, ly: float = 1.0, lz: float = 1.0) -> TetMesh:
    nn, ne = C.c_int(), C.c_int()
    _capi.call("pa_build_box", nx, ny, nz, lx, ly, lz, C.byref(nn), C.byref(ne), None, None)
    coords = np.empty((nn.value, 3), np.float64)
    conn = np.empty((ne.value, 4), np.int32)
    _capi.call("pa_build_box", nx, ny, nz, lx, ly, lz, C.byref(nn), C.byref(ne), _ptr(coords), _ptr(conn))
    return TetMesh(coords, conn)


def nodes_on_plane(mesh: TetMesh, axis: int, value: float, tol: float = 1e-12) -> np.ndarray:
    c = np.ascontiguousarray(mesh.coords)
    n = C.c_int()
    _capi.call("pa_nodes_on_plane", mesh.num_nodes, _ptr(c), axis, value, tol, None, C.byref(n))
    out = np.empty(n.value, np.int32)
    _capi.call("pa_nodes_on_plane", mesh.num_nodes, _ptr(c), axis, value, tol, _ptr(out), C.byref(n))
    return out


def faces_on_plane(mesh: TetMesh, axis: int, value: float, tol: float = 1e-12):
    c = np.ascontiguousarray(mesh.coords)
    e = np.ascontiguousarray(mesh.conn)
    n = C.c_int()
    _capi.call("pa_faces_on_plane", mesh.num_elems, _ptr(e), _ptr(c), axis, value, tol, None, None, C.byref(n))
    fv = np.empty((n.value, 3), np.int32)
    fe = np.empty(n.value, np.int32)
    _capi.call("pa_faces_on_plane", mesh.num_elems, _ptr(e), _ptr(c), axis, value, tol, _ptr(fv), _ptr(fe),
               C.byref(n))
    return fv, fe


def smoothed_density(nx: int, ny: int, nz: int, mesh: TetMesh, seed: int = 20260101, sigma: float = 2.0,
                     rho_min: float = 1e-3) -> np.ndarray:
    """Seeded nodal densities of the scale workloads: uniform noise on the node lattice,
    Gaussian filter (sigma in cells, reflect), standardised, sigmoid(2x) into [rho_min, 1].
    Returned in the mesh's node numbering."""
    lat = density_lattice(nx, ny, nz, seed, sigma, rho_min)
    c = mesh.coords
    return lattice_to_nodes(lat, c, c[:, 0].max() / nx, c[:, 1].max() / ny, c[:, 2].max() / nz)


def density_lattice(nx: int, ny: int, nz: int, seed: int = 20260101, sigma: float = 2.0,
                    rho_min: float = 1e-3) -> np.ndarray:
    """The density field of smoothed_density on the node lattice, indexed [k, j, i]."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    g = rng.uniform(0.0, 1.0, size=(nz + 1, ny + 1, nx + 1))
    g = gaussian_filter(g, sigma=sigma, mode="reflect")
    g = (g - g.mean()) / g.std()
    return rho_min + (1.0 - rho_min) / (1.0 + np.exp(-2.0 * g))


def lattice_to_nodes(lat: np.ndarray, coords: np.ndarray, hx: float, hy: float, hz: float) -> np.ndarray:
    """Values of a [k, j, i] lattice field at mesh nodes lying on the lattice of spacing
    (hx, hy, hz) from the origin, in the mesh's node numbering."""
    i = np.rint(coords[:, 0] / hx).astype(np.int64)
    j = np.rint(coords[:, 1] / hy).astype(np.int64)
    k = np.rint(coords[:, 2] / hz).astype(np.int64)
    return np.ascontiguousarray(lat[k, j, i])


def node_graph(mesh: TetMesh):
    """Block graph (rowmap, colidx) of the mesh, built on the host by the library."""
    conn = np.ascontiguousarray(mesh.conn, np.int32)
    rm = np.empty(mesh.num_nodes + 1, np.int32)
    nnz = C.c_int()
    _capi.call("pa_node_graph", mesh.num_nodes, mesh.num_elems, _ptr(conn), _ptr(rm), None, C.byref(nnz))
    ci = np.empty(nnz.value, np.int32)
    _capi.call("pa_node_graph", mesh.num_nodes, mesh.num_elems, _ptr(conn), _ptr(rm), _ptr(ci), C.byref(nnz))
    return rm, ci


def infer_box(mesh) -> tuple[int, int, int] | None:
    """(nx, ny, nz) when the nodes form a full tensor lattice with 6 tets per hex (the
    library still checks, when the box is declared, that the lattice is uniform and every
    hex carries the same six tets)."""
    X = mesh.coords
    n = [len(np.unique(X[:, d])) - 1 for d in range(3)]
    if min(n) < 1:
        return None
    if (n[0] + 1) * (n[1] + 1) * (n[2] + 1) != mesh.num_nodes or 6 * n[0] * n[1] * n[2] != mesh.num_elems:
        return None
    return (n[0], n[1], n[2])
