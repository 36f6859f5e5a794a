"""ctypes binding of ``include/plato_mi355x.h`` (libplato_mi355x.so, built in-tree).

The library is the product: there is no Python or CPU fallback for any compute entry
point.  Loading fails loudly when the shared object has not been built.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

from .build import LIB, ROOT

HEADER = ROOT / "include" / "plato_mi355x.h"

PA_OK = 0
STATUS_NAMES = {0: "PA_OK", 1: "PA_ERR_ARG", 2: "PA_ERR_HIP", 3: "PA_ERR_MESH", 4: "PA_ERR_STATE",
                5: "PA_ERR_NOT_CONVERGED", 6: "PA_ERR_UNSUPPORTED"}


PA_ERR_MESH = 3  # include/plato_mi355x.h: mesh/graph inconsistent


class PaError(RuntimeError):
    def __init__(self, status: int, where: str, msg: str):
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class SolverOpts(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("max_iters", C.c_int), ("check_every", C.c_int),
                ("precision", C.c_int), ("use_x0", C.c_int), ("inner_tolerance", C.c_double),
                ("preconditioner", C.c_int)]


class SolverStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("outer_iterations", C.c_int), ("converged", C.c_int),
                ("rel_residual", C.c_double)]


_P = C.c_void_p
_D = C.c_double
_I = C.c_int
_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int32)
_PIN = C.POINTER(C.c_int)

_SIGS = {
    "pa_last_error": (C.c_char_p, []),
    "pa_build_box": (_I, [_I, _I, _I, _D, _D, _D, _PIN, _PIN, _P, _P]),
    "pa_nodes_on_plane": (_I, [_I, _P, _I, _D, _D, _P, _PIN]),
    "pa_faces_on_plane": (_I, [_I, _P, _P, _I, _D, _D, _P, _P, _PIN]),
    "pa_node_graph": (_I, [_I, _I, _P, _P, _P, _PIN]),
    "pa_create": (_I, [C.POINTER(_P), _I, _I, _P, _P]),
    "pa_destroy": (_I, [_P]),
    "pa_set_material": (_I, [_P, _D, _D]),
    "pa_set_penalty": (_I, [_P, _D, _D]),
    "pa_add_essential_bc": (_I, [_P, _I, _P, _I, _D]),
    "pa_add_traction": (_I, [_P, _I, _P, _P, _P]),
    "pa_set_graph": (_I, [_P, _P, _P]),
    "pa_finalize": (_I, [_P]),
    "pa_set_stream": (_I, [_P, _P]),
    "pa_graph_size": (_I, [_P, _PIN, _PIN]),
    "pa_copy_graph": (_I, [_P, _P, _P, _I]),
    "pa_vf_value": (_I, [_P, _P, _P, _P]),
    "pa_vf_gradient_u": (_I, [_P, _P, _P, _P, _P]),
    "pa_vf_gradient_z": (_I, [_P, _P, _P, _P]),
    "pa_vf_gradient_x": (_I, [_P, _P, _P, _P]),
    "pa_sf_value": (_I, [_P, _P, _P, _PD]),
    "pa_sf_gradient_u": (_I, [_P, _P, _P, _P]),
    "pa_sf_gradient_z": (_I, [_P, _P, _P, _P, _PD]),
    "pa_sf_gradient_x": (_I, [_P, _P, _P, _P]),
    "pa_apply_constraints": (_I, [_P, _P, _P, _D]),
    "pa_solve": (_I, [_P, _P, _P, _P, C.POINTER(SolverOpts), C.POINTER(SolverStats)]),
    "pa_solve_noalloc": (_I, [_P, _P, _P, _P, C.POINTER(SolverOpts), C.POINTER(SolverStats)]),
    "pa_solution": (_I, [_P, _P, _P, _P, C.POINTER(SolverOpts), C.POINTER(SolverStats)]),
    "pa_solution_deflated": (_I, [_P, _P, _P, _P, _P, _I, C.POINTER(SolverOpts), C.POINTER(SolverStats)]),
    "pa_compliance_gradient": (_I, [_P, _P, _P, _P, _PD]),
    "pa_stress_pnorm": (_I, [_P, _P, _P, _I, _D, _D, _PD, _P, _P]),
    "pa_stress_pnorm_gradient_x": (_I, [_P, _P, _P, _I, _D, _D, _P, _PD]),
    "pa_dRdz_T_apply": (_I, [_P, _P, _P, _P, _P]),
    "pa_matvec_plus_vector": (_I, [_P, C.c_int, C.c_double, _P, _P, _P]),
    "pa_stress_pnorm_gradient": (_I, [_P, _P, _P, _I, _D, _D, _P, _P, _PD, C.POINTER(SolverOpts),
                                      C.POINTER(SolverStats)]),
    "pa_smooth": (_I, [_P, _P, _P, _I, _D]),
    "pa_set_box": (_I, [_P, _I, _I, _I]),
    "pa_box_assembly": (_I, [_P, _PIN]),
    "pa_set_penalty_model": (_I, [_P, _I, _D, _D, _D]),
    "pa_volume": (_I, [_P, _P, _I, _D, _D, _D, _PD, _P, _P]),
    "pa_unet_last_error": (C.c_char_p, []),
    "pa_unet_create": (_I, [C.POINTER(_P), _I, _I, _I, _I, C.c_ulonglong]),
    "pa_unet_destroy": (_I, [_P]),
    "pa_unet_set_stream": (_I, [_P, _P]),
    "pa_unet_num_params": (_I, [_P, C.POINTER(C.c_size_t)]),
    "pa_unet_params": (_I, [_P, _P, _P]),
    "pa_unet_forward": (_I, [_P, _P, _I, _P]),
    "pa_unet_gradient": (_I, [_P, _P, _P, _I, _P, _PD]),
    "pa_unet_train_step": (_I, [_P, _P, _P, _I, C.c_float, _PD]),
}

_lib = None


def header_symbols() -> list[str]:
    """Function names declared in include/plato_mi355x.h."""
    txt = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(pa_\w+)\s*\(", txt, re.M)))


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not Path(LIB).exists():
            raise ImportError(f"{LIB} is not built; run `python -m platoanalyze_amd.build` "
                              "(the HIP library is required, there is no fallback)")
        L = C.CDLL(str(LIB))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int, where: str) -> None:
    if status != PA_OK:
        msg = lib().pa_last_error().decode()
        raise PaError(status, where, msg)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)


def ucall(name: str, *args) -> None:
    """Call a pa_unet_* entry point (errors from pa_unet_last_error)."""
    status = getattr(lib(), name)(*args)
    if status != PA_OK:
        raise PaError(status, name, lib().pa_unet_last_error().decode())
