"""Loader of the in-tree HIP library ``libtdyhip.so`` (ctypes).

There is no fallback: if the library is missing or the GPU is absent, calls
raise.  ``load()`` only opens the shared object (it works without a GPU);
device memory is needed only for the evaluation calls.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtdyhip.so")

_lib = None

# every symbol declared in include/tdyhip.h
EXPORTS = [
    "tdyhip_last_error", "tdyhip_create", "tdyhip_destroy", "tdyhip_set_stream", "tdyhip_sizes",
    "tdyhip_jacobian_pattern", "tdyhip_jacobian_pattern_device", "tdyhip_ifunction", "tdyhip_ijacobian",
    "tdyhip_ifunction_ijacobian", "tdyhip_set_dtime", "tdyhip_snes_presolve", "tdyhip_snes_function",
    "tdyhip_snes_jacobian", "tdyhip_trans_nnz", "tdyhip_trans_csr", "tdyhip_gravdis",
    "tdyhip_solver_defaults", "tdyhip_newton_snes", "tdyhip_run_snes_steps", "tdyhip_run_ts_be",
    "tdyhip_run_tdyti", "tdyhip_set_pc_partition", "tdyhip_kernel_timers", "tdyhip_kernel_timer_report",
    "tdyhip_device_bytes", "tdyhip_diagnostics", "tdyhip_get_stream", "tdyhip_cancel",
    "tdyhip_setup_info", "tdyhip_setup_plan", "tdyhip_recover_velocity", "tdyhip_set_source_sink",
    "tdyhip_ifunction_noalloc", "tdyhip_batch_create", "tdyhip_batch_destroy", "tdyhip_batch_add",
    "tdyhip_batch_remove", "tdyhip_batch_size", "tdyhip_batch_round", "tdyhip_fuser_create",
    "tdyhip_fuser_destroy", "tdyhip_set_fuser", "tdyhip_fuser_stats", "tdyhip_fuser_set_window",
]


ERR_ARG = 1
ERR_SINGULAR = 3
ERR_NOCONV = 4
ERR_CANCELLED = 6


class TdyHipError(RuntimeError):
    """A nonzero TDYHIP_ERR_* return; ``code`` holds it."""

    def __init__(self, msg, code=0):
        super().__init__(msg)
        self.code = code


class Mesh(C.Structure):
    _fields_ = [("nc", C.c_int32), ("nf", C.c_int32), ("nv", C.c_int32), ("nbnd", C.c_int32),
                ("cell_verts", C.c_void_p), ("cell_faces", C.c_void_p), ("face_verts", C.c_void_p),
                ("face_cells", C.c_void_p), ("face_bc", C.c_void_p), ("vert_xyz", C.c_void_p),
                ("cell_cen", C.c_void_p), ("cell_vol", C.c_void_p), ("face_cen", C.c_void_p),
                ("face_area", C.c_void_p), ("face_normal", C.c_void_p)]


class Materials(C.Structure):
    _fields_ = [("vg_m", C.c_double), ("vg_alpha", C.c_double), ("poly_x0", C.c_double),
                ("poly_x1", C.c_double), ("poly_x2", C.c_double), ("poly_dx", C.c_double),
                ("gravity", C.c_double), ("pref", C.c_double), ("density_type", C.c_int32),
                ("gmatrix_method", C.c_int32), ("K", C.c_void_p), ("porosity", C.c_void_p),
                ("Sr", C.c_void_p), ("p_bnd_fn", C.c_void_p)]


class SolverOpts(C.Structure):
    _fields_ = [("snes_rtol", C.c_double), ("snes_atol", C.c_double), ("snes_stol", C.c_double),
                ("snes_max_it", C.c_int32), ("ksp_rtol", C.c_double), ("ksp_atol", C.c_double),
                ("ksp_max_it", C.c_int32), ("ksp_restart", C.c_int32), ("krylov_fp32", C.c_int32),
                ("pc_block", C.c_int32), ("hip_graphs", C.c_int32),
                ("pc_coarse", C.c_int32)]


class SolverStats(C.Structure):
    _fields_ = [("newton_its", C.c_int64), ("linear_its", C.c_int64), ("steps", C.c_int64),
                ("cuts", C.c_int64), ("status", C.c_int32), ("krylov_fallbacks", C.c_int32), ("last_fnorm", C.c_double)]


def load():
    """Open libtdyhip.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TdyHipError("libtdyhip.so is not built; run __graft_entry__.build() or python -m tdycore_amd.build")
    # torch ships its own HIP runtime under the unversioned name libamdhip64.so.
    # Loading torch first makes this library's libamdhip64.so.7 dependency
    # resolve to that already-loaded runtime (same SONAME), so the process has
    # one HIP runtime and device pointers from torch tensors are valid here.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    P, D, I, L64 = C.c_void_p, C.c_double, C.c_int, C.c_int64
    L.tdyhip_last_error.restype = C.c_char_p
    sig = {
        "tdyhip_create": [P, P, P], "tdyhip_destroy": [P], "tdyhip_set_stream": [P, P],
        "tdyhip_sizes": [P, P], "tdyhip_jacobian_pattern": [P, P, P],
        "tdyhip_jacobian_pattern_device": [P, P, P],
        "tdyhip_ifunction": [P, D, P, P, P, P], "tdyhip_ijacobian": [P, D, P, P, D, P, P],
        "tdyhip_ifunction_ijacobian": [P, D, P, P, D, P, P, P], "tdyhip_set_dtime": [P, D],
        "tdyhip_snes_presolve": [P, P, P], "tdyhip_snes_function": [P, P, P, P],
        "tdyhip_snes_jacobian": [P, P, P, P], "tdyhip_trans_nnz": [P, P],
        "tdyhip_trans_csr": [P, P, P, P], "tdyhip_gravdis": [P, P],
        "tdyhip_newton_snes": [P, P, P, P], "tdyhip_run_snes_steps": [P, P, D, C.c_int32, P, P],
        "tdyhip_run_ts_be": [P, P, D, C.c_int32, P, P], "tdyhip_run_tdyti": [P, P, D, D, D, D, P, P],
        "tdyhip_set_pc_partition": [P, P], "tdyhip_kernel_timers": [P, C.c_int32],
        "tdyhip_kernel_timer_report": [P, C.c_char_p, C.c_int64], "tdyhip_device_bytes": [P, P],
        "tdyhip_diagnostics": [P, P, P, P, P, P], "tdyhip_get_stream": [P, P], "tdyhip_cancel": [P, C.c_int32], "tdyhip_setup_info": [P, P, P], "tdyhip_setup_plan": [P, P, P],
        "tdyhip_recover_velocity": [P, P, P, P, P], "tdyhip_set_source_sink": [P, P],
        "tdyhip_ifunction_noalloc": [P, D, P, P, P, P],
        "tdyhip_batch_create": [P], "tdyhip_batch_destroy": [P],
        "tdyhip_batch_add": [P, P, P, P, D, D, D, D, P], "tdyhip_batch_remove": [P, C.c_int32],
        "tdyhip_batch_size": [P, P], "tdyhip_batch_round": [P, C.c_int32, P, P, P, P],
        "tdyhip_fuser_create": [P], "tdyhip_fuser_destroy": [P], "tdyhip_set_fuser": [P, P],
        "tdyhip_fuser_stats": [P, P], "tdyhip_fuser_set_window": [P, C.c_int32],
        # test hook (exported, not declared in the header)
        "tdyhip_test_dense_inverse": [P, C.c_int32, C.c_int32, P],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = I
    L.tdyhip_solver_defaults.argtypes = [P]
    L.tdyhip_solver_defaults.restype = None
    _lib = L
    return L


def check(rc: int):
    if rc != 0:
        msg = load().tdyhip_last_error().decode(errors="replace")
        raise TdyHipError(f"tdyhip error {rc}: {msg}", rc)
