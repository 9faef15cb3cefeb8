"""ctypes binding of libatv_hip.so (the C ABI declared in include/atvh.h).

There is no fallback: if the HIP library is missing or fails to load, every
entry point raises. Numpy arrays are passed as plain pointers.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libatv_hip.so")
INCLUDE_H = os.path.join(os.path.dirname(_PKG), "include", "atvh.h")

OK = 0
E_ARG, E_HIP, E_STATE, E_CAPACITY, E_NOMEM = -1, -2, -3, -4, -5
PREC_FP64, PREC_MIXED = 0, 1
PHASE_INITIAL, PHASE_FINAL = 0, 1
MEM_FRAME, MEM_NLIST, MEM_PAIR, MEM_PME, MEM_INTEGRATOR = 0, 1, 2, 3, 4

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_int64)
_vp = C.c_void_p

# name -> (restype, argtypes). Mirrors include/atvh.h one to one.
SIGNATURES = {
    "atvh_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "atvh_ctx_destroy": (C.c_int, [_vp]),
    "atvh_last_error": (C.c_char_p, [_vp]),
    "atvh_ctx_set_stream": (C.c_int, [_vp, _vp]),
    "atvh_ctx_stream": (_vp, [_vp]),
    "atvh_ctx_synchronize": (C.c_int, [_vp]),
    "atvh_ctx_enable_timing": (C.c_int, [_vp, C.c_int]),
    "atvh_ctx_timing_only": (C.c_int, [_vp, C.c_char_p]),
    "atvh_ctx_kernel_time": (C.c_int, [_vp, C.c_char_p, _dp, _lp]),
    "atvh_ctx_reset_timing": (C.c_int, [_vp]),
    "atvh_ctx_mem_info": (C.c_int, [_vp, _lp, _lp]),
    "atvh_system_create": (C.c_int, [_vp, C.c_int64, _dp, C.POINTER(_vp)]),
    "atvh_system_destroy": (C.c_int, [_vp]),
    "atvh_system_natoms": (C.c_int64, [_vp]),
    "atvh_frame_upload": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _dp]),
    "atvh_frame_download": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "atvh_system_box": (C.c_int, [_vp, _dp]),
    "atvh_system_memory": (C.c_int, [_vp, C.c_int, _lp]),
    "atvh_nlist_configure": (C.c_int, [_vp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int]),
    "atvh_nlist_build": (C.c_int, [_vp]),
    "atvh_nlist_decide_and_rebuild": (C.c_int, [_vp]),
    "atvh_nlist_brick_shape": (C.c_int, [_vp, _ip, _ip]),
    "atvh_nlist_set_brick_shape": (C.c_int, [_vp, C.c_int]),
    "atvh_nlist_brick_fill": (C.c_int, [_vp, _ip, _ip]),
    "atvh_nlist_set_brick_limit": (C.c_int, [_vp, C.c_int]),
    "atvh_nlist_autotune": (C.c_int, [_vp, C.c_int, C.c_int, _ip, _dp]),
    "atvh_nlist_set_capacity": (C.c_int, [_vp, C.c_int]),
    "atvh_nlist_stats": (C.c_int, [_vp, _lp, _lp, _ip, _ip]),
    "atvh_nlist_export_pairs": (C.c_int, [_vp, _lp, C.c_int64, _lp]),
    "atvh_pair_lj_configure": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_int, _dp, _dp]),
    "atvh_pair_ljcoul_configure": (C.c_int, [_vp, C.c_int, _dp, _dp, _ip, C.c_double, C.c_double, C.c_int]),
    "atvh_pair_compute": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "atvh_pme_configure": (C.c_int, [_vp, C.c_double, C.c_double, C.c_int]),
    "atvh_pme_compute": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "atvh_pme_info": (C.c_int, [_vp, _ip, _dp]),
    "atvh_energy": (C.c_int, [_vp, _dp, _dp, _dp]),
    "atvh_kinetic_stress": (C.c_int, [_vp]),
    "atvh_stress": (C.c_int, [_vp, _dp, _dp, _dp]),
    "atvh_integrator_configure": (C.c_int, [_vp, C.c_double, C.c_int, C.c_double, C.c_double, C.c_int64]),
    "atvh_step_vvnh": (C.c_int, [_vp, C.c_int]),
    "atvh_thermostat_state": (C.c_int, [_vp, _dp, _dp]),
    "atvh_npt_configure": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_int, C.c_int, _ip, C.c_double]),
    "atvh_npt_set_run_span": (C.c_int, [_vp, C.c_int64, C.c_int64]),
    "atvh_npt_thermo": (C.c_int, [_vp, _dp, _dp, C.c_int]),
    "atvh_npt_state": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp]),
    "atvh_md_run": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "atvh_md_thermo": (C.c_int, [_vp, _dp, _dp, C.c_int]),
    "atvh_md_run_noalloc": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp, _dp]),
    "atvh_batch_create": (C.c_int, [_vp, C.POINTER(_vp), C.c_int, C.POINTER(_vp)]),
    "atvh_batch_destroy": (C.c_int, [_vp]),
    "atvh_batch_set_graph": (C.c_int, [_vp, C.c_int]),
    "atvh_batch_run": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "atvh_batch_exchange": (C.c_int, [_vp, C.c_int64, C.c_uint64, _ip]),
    "atvh_batch_ladder": (C.c_int, [_vp, _ip, _dp]),
    "atvh_segment_bytes": (C.c_int64, [_vp]),
    "atvh_server_create": (C.c_int, [_vp, C.c_int64, C.c_int, C.c_int, C.POINTER(_vp)]),
    "atvh_server_destroy": (C.c_int, [_vp]),
    "atvh_server_submit": (C.c_int, [_vp, _vp, _lp]),
    "atvh_server_advance": (C.c_int, [_vp, _ip, _ip]),
    "atvh_server_poll": (C.c_int, [_vp, _lp, C.c_int, _ip]),
    "atvh_server_fetch": (C.c_int, [_vp, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "atvh_server_usage": (C.c_int, [_vp, _lp, _lp]),
    "atvh_util_stream_copy": (C.c_int, [_vp, C.c_int64, C.c_int, _dp]),
    "atvh_util_fcc_lattice": (C.c_int64, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int64, _dp, _dp, _dp]),
    "atvh_util_rocksalt_lattice": (C.c_int64, [C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _ip,
                                               _dp, C.c_double]),
    "atvh_util_init_velocity": (C.c_int, [C.c_int64, _dp, C.c_double, C.c_uint32, C.c_int64, _dp, _dp, _dp]),
}

_lib = None


class HipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"atvh error {code}: {msg}")
        self.code = code


def load():
    """Load libatv_hip.so; raise if it is absent (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built: run __graft_entry__.build() (hipcc, gfx950). "
                "The HIP backend has no CPU fallback.")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "float64 C-contiguous array required"
    return a.ctypes.data_as(_dp)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"], "int32 C-contiguous array required"
    return a.ctypes.data_as(_ip)


def lptr(a):
    if a is None:
        return None
    assert a.dtype == np.int64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_lp)
