"""MI355X-native Tet4 SIMP elastostatics path of Plato Analyze.

The compute lives in ``lib/libplato_mi355x.so`` (hand-written gfx950 HIP kernels behind
the C ABI of ``include/plato_mi355x.h``).  Python mirrors the reference's drop-in
surface (``Plato::Elliptic::VectorFunction``, ``PhysicsScalarFunction``, ``Problem``)
over torch device tensors; torch is used only for device memory and streams.
"""
from .mesh import TetMesh, build_box, faces_on_plane, nodes_on_plane, smoothed_density  # noqa: F401

__all__ = ["TetMesh", "build_box", "faces_on_plane", "nodes_on_plane", "smoothed_density"]
