"""ctypes loader for libmcsart_hip.so (fails loudly when the library is missing)."""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmcsart_hip.so")

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_SZ = ctypes.c_size_t

# Every exported symbol and its argument types; the header include/mcsart_hip.h is the source.
SIGNATURES = {
    "mcsart_status_string": ([_I], ctypes.c_char_p),
    "mcsart_last_error": ([], ctypes.c_char_p),
    "mcsart_version": ([], _I),
    "mcsart_device_count": ([ctypes.POINTER(_I)], _I),
    "mcsart_set_device": ([_I], _I),
    "mcsart_malloc": ([ctypes.POINTER(_P), _SZ], _I),
    "mcsart_free": ([_P], _I),
    "mcsart_memcpy_htod": ([_P, _P, _SZ], _I),
    "mcsart_memcpy_dtoh": ([_P, _P, _SZ], _I),
    "mcsart_memcpy_dtod": ([_P, _P, _SZ], _I),
    "mcsart_memset": ([_P, _I, _SZ], _I),
    "mcsart_device_synchronize": ([], _I),
    "mcsart_mem_info": ([ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)], _I),
    "mcsart_ctx_create": ([ctypes.POINTER(_P), _I], _I),
    "mcsart_ctx_destroy": ([_P], _I),
    "mcsart_ctx_synchronize": ([_P], _I),
    "mcsart_projection": ([_P, _P, _F, _F, _F, _F, _I, _F, _F, _I, _F, _I, _I, _I], _I),
    "mcsart_backprojection": ([_P, _P, _F, _F, _F, _F, _I, _F, _F, _I, _F, _I, _I, _I], _I),
    "mcsart_deform": ([_P, _P, _I, _I, _I, _F, _F] + [_P] * 6, _I),
    "mcsart_deform2": ([_P, _P, _I, _I, _I, _F, _F] + [_P] * 6, _I),
    "mcsart_update_udvf": ([_P] * 8 + [_F, _F, _I, _I, _I, _I], _I),
    "mcsart_add": ([_P, _P, _I, _I, _I, _F], _I),
    "mcsart_add2": ([_P, _P, _I, _I, _I, _P, _F, _I], _I),
    "mcsart_initial": ([_P, _I, _I, _I, _F], _I),
    "mcsart_initial2": ([_P, _I, _I, _I, _P, _F, _F], _I),
    "mcsart_division": ([_P, _P, _I, _I, _I], _I),
    "mcsart_diff": ([_P, _P, _I, _I, _I, _I], _I),
    "mcsart_diff_forward": ([_P, _P, _I, _I, _I, _I], _I),
    "mcsart_multiple": ([_P, _I, _I, _I, _P, _F, _I], _I),
    "mcsart_invert_dvf": ([_P] * 6 + [_I, _I, _I, _I], _I),
    "mcsart_generate_proj": ([_P] * 11 + [_I, _I, _F, _F, _F, _I, _F, _F, _I, _F, _I, _I, _I], _I),
    "mcsart_sart_residual": ([_P, _P, _P, _I, _P] + [_F, _F, _F, _F, _I, _F, _F, _I, _F, _I, _I, _I], _I),
    "mcsart_projection_noalloc": ([_P, _P, _P, _F, _F, _F, _F, _I, _F, _F, _I, _F, _I, _I, _I], _I),
    "mcsart_projection_noalloc_max_voxels": ([], _SZ),
    "mcsart_sart_backprojection": ([_P, _P, _F, _F, _F, _F, _I, _F, _F, _I, _F, _I, _I, _I], _I),
    "mcsart_motion_prep": ([_P] * 8 + [_F, _F, _I, _I, _I], _I),
    "mcsart_motion_update": ([_P] * 8 + [_F, _F, _I, _I, _I], _I),
    "mcsart_aumisart": ([_P] * 6 + [_I] * 8 + [_P] + [_F] * 8 + [_P] * 4, _I),
    "mcsart_aumisart_motion": ([_P] * 6 + [_I] * 8 + [_P] + [_F] * 8 + [_P] * 4 + [_P] * 6, _I),
    "mcsart_job_bytes": ([_P, _I, _I, _I], _SZ),
    "mcsart_job_create": ([_P, _P, _I, _P, _P, _P, _P, _I, _P, ctypes.POINTER(_P)], _I),
    "mcsart_job_set_projections": ([_P, _P], _I),
    "mcsart_job_set_projections_device": ([_P, _P], _I),
    "mcsart_job_set_image": ([_P, _P], _I),
    "mcsart_job_set_fields": ([_P, _P], _I),
    "mcsart_job_reset": ([_P], _I),
    "mcsart_job_run": ([_P], _I),
    "mcsart_job_run_sweep": ([_P, _I], _I),
    "mcsart_job_set_graphs": ([_P, _I], _I),
    "mcsart_job_set_culling": ([_P, _I], _I),
    "mcsart_view_voxels": ([_P, _I, _P, _P], _I),
    "mcsart_job_get_image": ([_P, _P], _I),
    "mcsart_job_get_norms": ([_P, _P], _I),
    "mcsart_job_get_residual": ([_P, _P], _I),
    "mcsart_job_get_fields": ([_P, _P], _I),
    "mcsart_job_device_image": ([_P, ctypes.POINTER(_P)], _I),
    "mcsart_job_destroy": ([_P], _I),
    "mcsart_job_set_tuning": ([_P, _P], _I),
    "mcsart_job_get_tuning": ([_P, _P], _I),
    "mcsart_job_autotune": ([_P, _P], _I),
    "mcsart_job_kernel_time": ([_P, _I, _I, ctypes.POINTER(ctypes.c_double)], _I),
    "mcsart_job_view_layouts": ([_P, _P], _I),
    "mcsart_ctx_enable_timing": ([_P, _I], _I),
    "mcsart_ctx_kernel_time": ([_P, _I, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_longlong)], _I),
    "mcsart_ctx_reset_timing": ([_P], _I),
    "mcsart_server_create": ([_I, _P, ctypes.POINTER(_P)], _I),
    "mcsart_server_submit": ([_P, _P, ctypes.POINTER(ctypes.c_longlong)], _I),
    "mcsart_server_wait": ([_P, ctypes.c_longlong, ctypes.c_double], _I),
    "mcsart_server_cancel": ([_P, ctypes.c_longlong], _I),
    "mcsart_server_drain": ([_P], _I),
    "mcsart_server_info": ([_P, ctypes.c_longlong, _P], _I),
    "mcsart_server_release": ([_P, ctypes.c_longlong], _I),
    "mcsart_server_memory": ([_P, ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)], _I),
    "mcsart_server_destroy": ([_P], _I),
    "mcsart_parse_geo": ([_I, _P, _P, _P], _I),
    "mcsart_parse_iter": ([_P, _P, _P], _I),
}
# Asynchronous operators: a context, then the synchronous operator's arguments.
ASYNC_OPS = ("projection", "backprojection", "deform", "deform2", "update_udvf", "add", "add2",
             "initial", "initial2", "division", "diff", "diff_forward", "sart_residual",
             "sart_backprojection", "motion_prep", "motion_update", "multiple", "invert_dvf",
             "generate_proj")
for _op in ASYNC_OPS:
    SIGNATURES[f"mcsart_{_op}_async"] = ([_P] + SIGNATURES[f"mcsart_{_op}"][0], _I)

_lib = None


class McsartError(RuntimeError):
    def __init__(self, status: int, where: str):
        lib = load()
        msg = lib.mcsart_last_error().decode(errors="replace")
        name = lib.mcsart_status_string(status).decode()
        super().__init__(f"{where}: {name} ({status}): {msg}")
        self.status = status


def load():
    """Load libmcsart_hip.so; raises if it has not been built (no fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m mcsart_amd.build`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


def check(status: int, where: str) -> None:
    if status != 0:
        raise McsartError(status, where)
