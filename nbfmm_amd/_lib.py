"""ctypes binding of libnbfmm_amd.so (include/nbfmm_amd.h)."""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libnbfmm_amd.so")
_lib = None

c_int, c_float, c_void_p, c_size_t, c_double, c_uint = (ctypes.c_int, ctypes.c_float, ctypes.c_void_p,
                                                        ctypes.c_size_t, ctypes.c_double, ctypes.c_uint)
P = c_void_p

# name -> argument types (every function returns int unless listed in _RESTYPE)
SIGNATURES = {
    "nbfmm_version": [],
    "nbfmm_last_error": [],
    "nbfmm_solver_create": [c_int, c_int, ctypes.POINTER(c_float), c_float, ctypes.POINTER(c_void_p)],
    "nbfmm_solver_destroy": [P],
    "nbfmm_solver_info": [P, ctypes.POINTER(c_int), ctypes.POINTER(c_int), ctypes.POINTER(c_int)],
    "nbfmm_solve": [P, c_int, P, P, P, P],
    "nbfmm_solve_noalloc": [c_int, c_int, ctypes.POINTER(c_float), c_float, P, P, P, P],
    "nbfmm_predo": [P, c_int, P, P, P],
    "nbfmm_p2p": [P, c_int, P],
    "nbfmm_p2m": [P, c_int, P],
    "nbfmm_m2m": [P, P],
    "nbfmm_m2l": [P, P],
    "nbfmm_l2l": [P, P],
    "nbfmm_l2p": [P, c_int, P],
    "nbfmm_postdo": [P, c_int, P, P],
    "nbfmm_buffer_bytes": [P, c_int, c_int, ctypes.POINTER(c_size_t)],
    "nbfmm_buffer_read": [P, c_int, c_int, P, P],
    "nbfmm_buffer_write": [P, c_int, c_int, P, P],
    "nbfmm_solver_flags": [P, ctypes.POINTER(c_uint), c_int],
    "nbfmm_update": [c_int, c_float, P, P, P, P],
    "nbfmm_step": [P, c_int, c_float, P, P, P, P, P],
    "nbfmm_run": [P, c_int, c_float, c_int, P, P, P, P, c_int, P],
    "nbfmm_profile_enable": [P, c_int],
    "nbfmm_p2p_time": [P, ctypes.POINTER(c_double), ctypes.POINTER(c_int), c_int],
    "nbfmm_stage_times": [P, ctypes.POINTER(c_double), ctypes.POINTER(c_int), c_int],
    "nbfmm_pair_count": [P, c_int, ctypes.POINTER(c_double), ctypes.POINTER(c_int)],
    "nbfmm_batch_create": [c_int, P, P, P, P, ctypes.POINTER(c_void_p)],
    "nbfmm_batch_destroy": [P],
    "nbfmm_batch_info": [P, ctypes.POINTER(c_int), ctypes.POINTER(c_int)],
    "nbfmm_batch_offsets": [P, P],
    "nbfmm_batch_set_ticks": [P, P],
    "nbfmm_batch_solve": [P, P, P, P, P],
    "nbfmm_batch_step": [P, P, P, P, P, P],
    "nbfmm_batch_run": [P, c_int, P, P, P, P, c_int, P],
    "nbfmm_batch_flags": [P, P, c_int],
    "nbfmm_batch_set_p2p_config": [P, c_int, c_int],
    "nbfmm_batch_set_warm_start": [P, c_int],
    "nbfmm_batch_balance": [P, P, P, P],
    "nbfmm_batch_get_p2p_config": [P, ctypes.POINTER(c_int), ctypes.POINTER(c_int)],
    "nbfmm_p2p_variant_count": [],
    "nbfmm_p2p_variant_info": [c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int), ctypes.POINTER(c_int)],
    "nbfmm_set_p2p_config": [P, c_int, c_int],
    "nbfmm_set_warm_start": [P, c_int],
    "nbfmm_set_small_front": [P, c_int],
    "nbfmm_prune": [P, c_int, P, P, P, ctypes.POINTER(c_int), P],
    "nbfmm_model_disk": [c_int, c_float, c_float, c_float, c_float, c_float, c_float, ctypes.c_uint, P, P, P, P],
    "nbfmm_model_disk_static": [c_int, c_float, c_float, c_float, c_float, ctypes.c_uint, P, P, P, P],
    "nbfmm_model_disk_center": [c_int, c_float, c_float, c_float, c_float, c_float, c_float, c_float, ctypes.c_uint,
                                P, P, P, P],
    "nbfmm_model_circle_center": [c_int, c_float, c_float, c_float, c_float, c_float, c_float, c_float, c_int,
                                  ctypes.c_uint, P, P, P, P],
    "nbfmm_model_rectangle": [c_int, c_float, c_float, c_float, c_float, c_float, ctypes.c_uint, P, P, P, P],
    "nbfmm_model_double_disk": [c_int, c_int, P, P, c_float, c_float, c_float, c_float, c_float, c_float, c_float,
                                c_float, c_int, ctypes.c_uint, P, P, P, P],
    "nbfmm_display": [c_int, c_int, c_int, c_float, ctypes.POINTER(c_float), P, P, P, P],
    "nbfmm_y4m_open": [ctypes.c_char_p, c_int, c_int, c_int, ctypes.POINTER(c_void_p)],
    "nbfmm_y4m_write_frame": [P, P, P],
    "nbfmm_y4m_close": [P, ctypes.POINTER(ctypes.c_long)],
    "nbfmm_get_p2p_config": [P, ctypes.POINTER(c_int), ctypes.POINTER(c_int)],
    "nbfmm_set_p2p_wait_polls": [P, c_int],
    "nbfmm_batch_set_p2p_wait_polls": [P, c_int],
    "nbfmm_autotune": [P, c_int, P, P, c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int),
                       ctypes.POINTER(c_float), P],
    "nbfmm_batch_solve_adjoint": [P, P, P, P, P, P, P],
    "nbfmm_batch_step_adjoint": [P, P, P, P, P, P, P],
    "nbfmm_batch_keep_sort": [P, c_int, P],
    "nbfmm_batch_step_adjoint_kept": [P, c_int, P, P, P, P],
    "nbfmm_batch_run_keep": [P, c_int, c_int, P, P, P, P, P],
    "nbfmm_solve_adjoint": [P, c_int, P, P, P, P, P, P],
    "nbfmm_step_adjoint": [P, c_int, c_float, P, P, P, P, P, P],
    "nbfmm_trajectory_workspace": [c_int, c_int, c_int, ctypes.POINTER(c_size_t), ctypes.POINTER(c_size_t)],
    "nbfmm_trajectory_forward": [P, c_int, c_float, c_int, c_int, P, P, P, P, P, c_int, P, P, P],
    "nbfmm_trajectory_backward": [P, c_int, c_float, c_int, c_int, P, P, c_int, P, P, P, P, P, P, P],
}
_RESTYPE = {"nbfmm_last_error": ctypes.c_char_p}


class NbfmmError(RuntimeError):
    CODES = {1: "invalid argument", 2: "out of memory", 3: "HIP error", 4: "capacity exceeded", 5: "no device"}

    def __init__(self, code: int, msg: str):
        self.code = code
        super().__init__(f"nbfmm error {code} ({self.CODES.get(code, '?')}): {msg}")


def lib_path() -> str:
    return _LIB_PATH


def lib():
    """Load the HIP library; raises if it was not built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise ImportError(f"{_LIB_PATH} is missing: build it with `python -m nbfmm_amd.build`")
        L = ctypes.CDLL(_LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPE.get(name, c_int)
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().nbfmm_last_error().decode(errors="replace")
        raise NbfmmError(rc, msg)
