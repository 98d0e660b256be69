"""ctypes binding of libstreams_mi355x.so (the C ABI of include/streams_mi355x.h).

The library is built in place next to this package (``make -C streams-2_amd``
or ``__graft_entry__.build()``).  Loading fails loudly when it is missing:
there is no CPU fallback of the solver path.
"""
from __future__ import annotations

import ctypes as C
import os

from .params import StParams, StScalars

ST_OK = 0
ST_ERR_ARG = 1
ST_ERR_HIP = 2
ST_ERR_NOMEM = 3
ST_ERR_STATE = 4
ST_ERR_NAN = 5

KERNEL_FAMILIES = ("aux", "euler_x", "visc", "euler_y", "euler_z", "sensor", "visc_div", "finish",
                   "update", "force_var", "dt", "step_end", "bc", "euler_xy", "weno_fix", "euler_zv")

LIB_NAME = "libstreams_mi355x.so"
_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_DIR, LIB_NAME)

_LIB = None

_SIGNATURES = {
    "st_solver_create": (C.c_int, [C.POINTER(StParams), C.POINTER(C.c_void_p)]),
    "st_solver_destroy": (C.c_int, [C.c_void_p]),
    "st_solver_footprint": (C.c_int64, [C.POINTER(StParams)]),
    "st_set_grid": (C.c_int, [C.c_void_p] + [C.POINTER(C.c_double)] * 13),
    "st_copy_from_field": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "st_copy_to_field": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "st_get_aux": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    "st_set_aux": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    "st_get_fl": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "st_get_scalars": (C.c_int, [C.c_void_p, C.POINTER(StScalars)]),
    "st_set_time": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_int64]),
    "st_update_ghost": (C.c_int, [C.c_void_p, C.c_int32]),
    "st_bcswap": (C.c_int, [C.c_void_p]),
    "st_compute_aux": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "st_visflx": (C.c_int, [C.c_void_p, C.c_int32]),
    "st_euler_x": (C.c_int, [C.c_void_p] + [C.c_int32] * 5),
    "st_euler_y": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "st_euler_z": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "st_get_eul_range": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "st_init_flux": (C.c_int, [C.c_void_p, C.c_double]),
    "st_update_flux": (C.c_int, [C.c_void_p, C.c_double]),
    "st_update_field": (C.c_int, [C.c_void_p]),
    "st_force_rhs": (C.c_int, [C.c_void_p]),
    "st_force_var": (C.c_int, [C.c_void_p]),
    "st_reset_forcing": (C.c_int, [C.c_void_p]),
    "st_bcextr_var": (C.c_int, [C.c_void_p, C.c_int32]),
    "st_bcswap_var": (C.c_int, [C.c_void_p, C.c_int32]),
    "st_recyc_exchange": (C.c_int, [C.c_void_p]),
    "st_zero_flux": (C.c_int, [C.c_void_p]),
    "st_bc_nr": (C.c_int, [C.c_void_p]),
    "st_set_recycling": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_double, C.POINTER(C.c_double),
                                   C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
                         + [C.POINTER(C.c_double)] * 4),
    "st_set_wmean": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "st_init_run": (C.c_int, [C.c_void_p]),
    "st_compute_dt": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "st_compute_dt_noalloc": (C.c_int, [C.POINTER(StParams), C.c_void_p, C.c_void_p, C.c_void_p]
                              + [C.c_void_p] * 7),
    "st_rk_sync": (C.c_int, [C.c_void_p]),
    "st_advance": (C.c_int, [C.c_void_p, C.c_int32]),
    "st_compute_residual": (C.c_int, [C.c_void_p, C.POINTER(StScalars)]),
    "st_count_weno": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "st_compute_stats": (C.c_int, [C.c_void_p, C.c_int32]),
    "st_get_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "st_set_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "st_kernel_times": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
    "st_set_tuning": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "st_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "st_synchronize": (C.c_int, [C.c_void_p]),
    "st_last_error": (C.c_char_p, []),
    "st_version": (C.c_char_p, []),
    "st_server_create": (C.c_int, [C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "st_server_destroy": (C.c_int, [C.c_void_p]),
    "st_server_submit": (C.c_int, [C.c_void_p, C.POINTER(StParams)] + [C.POINTER(C.c_double)] * 14
                         + [C.c_int32, C.POINTER(C.c_int64)]),
    "st_server_submit2": (C.c_int, [C.c_void_p, C.POINTER(StParams)] + [C.POINTER(C.c_double)] * 14
                          + [C.c_int32, C.c_double, C.c_int64, C.c_int32, C.c_char_p, C.POINTER(C.c_int64)]),
    "st_server_wait": (C.c_int, [C.c_void_p, C.c_int64, C.c_double, C.POINTER(C.c_double), C.POINTER(StScalars),
                                 C.POINTER(C.c_double)]),
    "st_server_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "st_server_set_batching": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64]),
    "st_server_batch_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "st_server_last_error": (C.c_char_p, []),
}


class StreamsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[st error {code}] {msg}")
        self.code = code


def lib(path: str = None):
    """Load the library once; raise if it has not been built.  The
    STREAMS_MI355X_LIB environment variable selects another build of the
    same library (A/B timing of kernel variants)."""
    global _LIB
    if path is None:
        path = os.environ.get("STREAMS_MI355X_LIB", LIB_PATH)
    if _LIB is None:
        if not os.path.exists(path):
            raise RuntimeError(f"{LIB_NAME} not built at {path}: run make -C streams-2_amd")
        L = C.CDLL(path)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def exported_symbols():
    return sorted(_SIGNATURES)


def check(rc: int):
    if rc != ST_OK:
        raise StreamsError(rc, lib().st_last_error().decode())
    return rc


def check_server(rc: int):
    if rc != ST_OK:
        raise StreamsError(rc, lib().st_server_last_error().decode())
    return rc
