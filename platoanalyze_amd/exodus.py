"""Exodus II mesh ingest (and a writer for tests and round trips) for Tet4 meshes with
named node sets and side sets, the form the reference's decks take their meshes in
(src/PlatoDriver.hpp:70-90 reads the "Input Mesh" through Omega_h; node sets name the
"Essential Boundary Conditions", side sets the "Natural Boundary Conditions").

Files are NetCDF "classic" (CDF-1) or "64-bit offset" (CDF-2), read with
scipy.io.netcdf_file.  NetCDF-4 (HDF5-based) Exodus files need an HDF5 reader, which this
environment does not have: `read_exodus` raises ValueError for them with that reason.

Exodus conventions used (Exodus II format: 1-based ids, per block `connect<k>` with
`elem_type`, coordinates `coord` or `coordx/coordy/coordz`, node sets `node_ns<k>`, side
sets `elem_ss<k>` / `side_ss<k>`, names in `eb_names` / `ns_names` / `ss_names`); TETRA4
side numbering: side 1 = nodes (1, 2, 4), 2 = (2, 3, 4), 3 = (1, 4, 3), 4 = (1, 3, 2).
"""
from __future__ import annotations

import dataclasses
from pathlib import Path

import numpy as np

from .mesh import TetMesh

TET_SIDES = ((0, 1, 3), (1, 2, 3), (0, 3, 2), (0, 2, 1))  # 0-based local nodes per side
_TET_TYPES = ("TETRA", "TETRA4", "TET4", "TET")
_LEN_NAME = 33


@dataclasses.dataclass
class ExodusMesh:
    mesh: TetMesh
    node_sets: dict        # name -> node ids (0-based int32)
    side_sets: dict        # name -> (face vertices [k,3] int32, owning elements [k] int32)
    blocks: dict           # name -> (first element, element count)


def _names(f, var: str, count: int, prefix: str, ids) -> list[str]:
    out = []
    if var in f.variables:
        raw = f.variables[var][:]
        for k in range(count):
            s = raw[k].tobytes().split(b"\0", 1)[0].decode("ascii", "replace").strip()
            out.append(s)
    else:
        out = [""] * count
    return [n if n else f"{prefix}{ids[k]}" for k, n in enumerate(out)]


def _dim(f, name: str) -> int:
    d = f.dimensions.get(name)
    return int(d) if d is not None else 0


def read_exodus(path: str | Path) -> ExodusMesh:
    """Read a Tet4 Exodus II file (all element blocks must be TETRA4)."""
    from scipy.io import netcdf_file

    path = Path(path)
    with open(path, "rb") as fh:
        magic = fh.read(4)
    if magic[:3] != b"CDF":
        raise ValueError(f"{path}: not a NetCDF classic / 64-bit-offset file (NetCDF-4/HDF5 Exodus "
                         "files need an HDF5 reader, which is not available here)")
    f = netcdf_file(str(path), "r", mmap=False)
    try:
        nn = _dim(f, "num_nodes")
        if "coord" in f.variables:
            X = np.array(f.variables["coord"][:], dtype=np.float64).T.copy()
        else:
            X = np.stack([np.array(f.variables[v][:], dtype=np.float64) for v in ("coordx", "coordy", "coordz")],
                         axis=1)
        if X.shape != (nn, 3):
            raise ValueError(f"{path}: expected 3-D coordinates of {nn} nodes, got {X.shape}")
        nblk = _dim(f, "num_el_blk")
        blk_ids = (np.array(f.variables["eb_prop1"][:]).tolist() if "eb_prop1" in f.variables
                   else list(range(1, nblk + 1)))
        blk_names = _names(f, "eb_names", nblk, "block_", blk_ids)
        conns, blocks, first = [], {}, 0
        for k in range(1, nblk + 1):
            var = f.variables[f"connect{k}"]
            etype = var.elem_type.decode() if isinstance(var.elem_type, bytes) else str(var.elem_type)
            if etype.upper() not in _TET_TYPES:
                raise ValueError(f"{path}: block {blk_names[k - 1]} has element type {etype}, only TETRA4 is supported")
            c = np.array(var[:], dtype=np.int64)
            if c.ndim != 2 or c.shape[1] != 4:
                raise ValueError(f"{path}: block {blk_names[k - 1]} is not 4-node")
            conns.append(c - 1)
            blocks[blk_names[k - 1]] = (first, c.shape[0])
            first += c.shape[0]
        conn = np.concatenate(conns).astype(np.int32) if conns else np.zeros((0, 4), np.int32)
        if conn.size and (conn.min() < 0 or conn.max() >= nn):
            raise ValueError(f"{path}: connectivity out of range")
        nns = _dim(f, "num_node_sets")
        ns_ids = (np.array(f.variables["ns_prop1"][:]).tolist() if "ns_prop1" in f.variables
                  else list(range(1, nns + 1)))
        node_sets = {}
        for k, name in enumerate(_names(f, "ns_names", nns, "nodeset_", ns_ids), start=1):
            node_sets[name] = np.array(f.variables[f"node_ns{k}"][:], dtype=np.int64).astype(np.int32) - 1
        nss = _dim(f, "num_side_sets")
        ss_ids = (np.array(f.variables["ss_prop1"][:]).tolist() if "ss_prop1" in f.variables
                  else list(range(1, nss + 1)))
        side_sets = {}
        for k, name in enumerate(_names(f, "ss_names", nss, "sideset_", ss_ids), start=1):
            el = np.array(f.variables[f"elem_ss{k}"][:], dtype=np.int64) - 1
            sd = np.array(f.variables[f"side_ss{k}"][:], dtype=np.int64) - 1
            if el.size and (el.min() < 0 or el.max() >= conn.shape[0] or sd.min() < 0 or sd.max() > 3):
                raise ValueError(f"{path}: side set {name} out of range")
            loc = np.array(TET_SIDES, dtype=np.int64)[sd]                       # [k, 3]
            fv = np.take_along_axis(conn[el].astype(np.int64), loc, axis=1)     # [k, 3]
            side_sets[name] = (fv.astype(np.int32), el.astype(np.int32))
    finally:
        f.close()
    return ExodusMesh(TetMesh(X, conn), node_sets, side_sets, blocks)


def write_exodus(path: str | Path, mesh: TetMesh, node_sets: dict | None = None, side_sets: dict | None = None,
                 block_name: str = "body") -> None:
    """Write a single-block TETRA4 Exodus II file (64-bit-offset NetCDF).  side_sets maps a
    name to (elements, local side index 0..3) pairs in the TETRA4 numbering above."""
    from scipy.io import netcdf_file

    node_sets = node_sets or {}
    side_sets = side_sets or {}
    f = netcdf_file(str(path), "w", version=2)
    try:
        f.title = b"platoanalyze_amd"
        f.api_version = np.float32(8.0)
        f.version = np.float32(8.0)
        f.floating_point_word_size = np.int32(8)
        f.file_size = np.int32(1)
        f.createDimension("time_step", None)  # the record dimension comes first
        f.createDimension("len_string", 33)
        f.createDimension("len_name", _LEN_NAME)
        f.createDimension("four", 4)
        f.createDimension("num_dim", 3)
        f.createDimension("num_nodes", mesh.num_nodes)
        f.createDimension("num_elem", mesh.num_elems)
        f.createDimension("num_el_blk", 1)
        t = f.createVariable("time_whole", "d", ("time_step",))
        del t
        c = f.createVariable("coord", "d", ("num_dim", "num_nodes"))
        c[:] = np.ascontiguousarray(mesh.coords.T)
        f.createDimension("num_el_in_blk1", mesh.num_elems)
        f.createDimension("num_nod_per_el1", 4)
        ebp = f.createVariable("eb_prop1", "i", ("num_el_blk",))
        ebp[:] = [1]
        ebp.name = b"ID"
        ebs = f.createVariable("eb_status", "i", ("num_el_blk",))
        ebs[:] = [1]
        _put_names(f, "eb_names", "num_el_blk", [block_name])
        con = f.createVariable("connect1", "i", ("num_el_in_blk1", "num_nod_per_el1"))
        con[:] = mesh.conn + 1
        con.elem_type = b"TETRA4"
        if node_sets:
            f.createDimension("num_node_sets", len(node_sets))
            p = f.createVariable("ns_prop1", "i", ("num_node_sets",))
            p[:] = np.arange(1, len(node_sets) + 1)
            p.name = b"ID"
            st = f.createVariable("ns_status", "i", ("num_node_sets",))
            st[:] = 1
            _put_names(f, "ns_names", "num_node_sets", list(node_sets))
            for k, (name, nodes) in enumerate(node_sets.items(), start=1):
                f.createDimension(f"num_nod_ns{k}", len(nodes))
                v = f.createVariable(f"node_ns{k}", "i", (f"num_nod_ns{k}",))
                v[:] = np.asarray(nodes) + 1
        if side_sets:
            f.createDimension("num_side_sets", len(side_sets))
            p = f.createVariable("ss_prop1", "i", ("num_side_sets",))
            p[:] = np.arange(1, len(side_sets) + 1)
            p.name = b"ID"
            st = f.createVariable("ss_status", "i", ("num_side_sets",))
            st[:] = 1
            _put_names(f, "ss_names", "num_side_sets", list(side_sets))
            for k, (name, (elems, sides)) in enumerate(side_sets.items(), start=1):
                f.createDimension(f"num_side_ss{k}", len(elems))
                e = f.createVariable(f"elem_ss{k}", "i", (f"num_side_ss{k}",))
                e[:] = np.asarray(elems) + 1
                s = f.createVariable(f"side_ss{k}", "i", (f"num_side_ss{k}",))
                s[:] = np.asarray(sides) + 1
    finally:
        f.close()


def _put_names(f, var: str, dim: str, names: list[str]) -> None:
    v = f.createVariable(var, "c", (dim, "len_name"))
    arr = np.zeros((len(names), _LEN_NAME), dtype="S1")
    for k, n in enumerate(names):
        b = n.encode("ascii")[:_LEN_NAME - 1]
        arr[k, :len(b)] = np.frombuffer(b, dtype="S1")
    v[:] = arr


def sides_of_faces(mesh: TetMesh, face_verts: np.ndarray, face_elems: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """(elements, TETRA4 local side index) of boundary faces given by their vertices."""
    sides = np.empty(len(face_elems), dtype=np.int32)
    for k, (fv, e) in enumerate(zip(face_verts, face_elems)):
        local = set(int(np.flatnonzero(mesh.conn[e] == v)[0]) for v in fv)
        for sidx, loc in enumerate(TET_SIDES):
            if set(loc) == local:
                sides[k] = sidx
                break
        else:
            raise ValueError(f"face {k} is not a side of element {e}")
    return np.asarray(face_elems, dtype=np.int32), sides


def model_from_exodus(ex: ExodusMesh, E: float = 1.0, nu: float = 0.3, penalty_exponent: float = 3.0,
                      penalty_minimum: float = 1e-9, essential: dict | None = None, tractions: dict | None = None,
                      **kw):
    """ElasticModel from named sets: essential = {node set: [(dof, value), ...]},
    tractions = {side set: (tx, ty, tz)}.  A mesh whose nodes form a uniform box lattice
    is declared a box (lattice assembly, multigrid available) unless `box` is given;
    box=None declares none."""
    from .elliptic import ElasticModel, EssentialBC, Traction
    from .mesh import infer_box

    bcs = []
    for name, dofs in (essential or {}).items():
        if name not in ex.node_sets:
            raise KeyError(f"node set {name!r} not in the mesh ({sorted(ex.node_sets)})")
        for dof, value in dofs:
            bcs.append(EssentialBC(ex.node_sets[name], int(dof), float(value)))
    trs = []
    for name, t in (tractions or {}).items():
        if name not in ex.side_sets:
            raise KeyError(f"side set {name!r} not in the mesh ({sorted(ex.side_sets)})")
        fv, fe = ex.side_sets[name]
        trs.append(Traction(fv, fe, tuple(float(v) for v in t)))
    box = kw.pop("box", "auto")
    if box != "auto":
        return ElasticModel(ex.mesh, E, nu, penalty_exponent, penalty_minimum, bcs, trs, box=box, **kw)
    # a tensor lattice with unequal spacings is no box: the handle then keeps the general assembly
    return ElasticModel(ex.mesh, E, nu, penalty_exponent, penalty_minimum, bcs, trs, box=infer_box(ex.mesh),
                        box_optional=True, **kw)
