"""ctypes binding of libfc3.so (the C ABI declared in include/fc3.h).

This is the Python counterpart of the Julia `ccall` layer: it loads the in-tree
library and declares every entry point's signature.  There is no fallback: if the
library is missing or fails to load, every call raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfc3.so")

FC3_OK = 0
FC3_ERR_ARGUMENT = 1
FC3_ERR_CUTOFF = 2
FC3_ERR_BLOCK_SIZE = 3
FC3_ERR_OUT_OF_MEMORY = 4
FC3_ERR_HIP = 5
FC3_ERR_UNSUPPORTED = 6
FC3_ERR_CAPACITY = 7

FC3_PREC_FP32 = 0
FC3_PREC_BF16X3 = 1

_c_int = ctypes.c_int
_c_double = ctypes.c_double
_c_size = ctypes.c_size_t
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_handle = ctypes.c_void_p

# name -> (restype, argtypes); kept in the order of include/fc3.h
class Snapshot(ctypes.Structure):
    """fc3_snapshot of fc3.h"""
    _fields_ = [("positions", ctypes.POINTER(ctypes.c_double)), ("n_atoms", ctypes.c_int),
                ("box", ctypes.c_double * 3), ("masses", ctypes.POINTER(ctypes.c_double)),
                ("phi", ctypes.c_void_p), ("k3", ctypes.c_void_p),
                ("freqs_sq", ctypes.POINTER(ctypes.c_double))]


FC3_BATCH_AUTOTUNE = 1
FC3_BATCH_NO_GRAPH = 2
FC3_BATCH_DEVICE_MODES = 4

SIGNATURES = {
    "fc3_version": (_c_int, []),
    "fc3_last_error": (ctypes.c_char_p, []),
    "fc3_status_string": (ctypes.c_char_p, [_c_int]),
    "fc3_create": (_c_int, [ctypes.POINTER(_handle), _c_int]),
    "fc3_destroy": (_c_int, [_handle]),
    "fc3_release_workspace": (_c_int, [_handle]),
    "fc3_synchronize": (_c_int, [_handle]),
    "fc3_get_device": (_c_int, [_handle, ctypes.POINTER(_c_int)]),
    "fc3_device_alloc": (_c_int, [_handle, ctypes.POINTER(_vp), _c_size]),
    "fc3_device_free": (_c_int, [_handle, _vp]),
    "fc3_memcpy_h2d": (_c_int, [_handle, _vp, _vp, _c_size]),
    "fc3_memcpy_d2h": (_c_int, [_handle, _vp, _vp, _c_size]),
    "fc3_memset_zero": (_c_int, [_handle, _vp, _c_size]),
    "fc3_device_memory": (_c_int, [_handle, ctypes.POINTER(_c_size), ctypes.POINTER(_c_size)]),
    "fc3_timer_start": (_c_int, [_handle]),
    "fc3_timer_stop": (_c_int, [_handle, ctypes.POINTER(ctypes.c_float)]),
    "fc3_profile_enable": (_c_int, [_handle, _c_int]),
    "fc3_profile_query": (_c_int, [_handle, _c_int, ctypes.POINTER(_c_double),
                                   ctypes.POINTER(_c_int)]),
    "fc3_second_order_lj": (_c_int, [_handle, _dp, _c_int, _dp, _c_double, _c_double, _c_double,
                                     _c_double, _c_double, _vp]),
    "fc3_second_order_sw": (_c_int, [_handle, _dp, _c_int, _dp, _dp, _c_double, _c_double, _vp]),
    "fc3_third_order_lj": (_c_int, [_handle, _dp, _c_int, _dp, _c_double, _c_double, _c_double,
                                    _c_double, _c_double, _dp, _vp, _vp]),
    "fc3_third_order_lj_noalloc": (_c_int, [_handle, _dp, _c_int, _dp, _c_double, _c_double,
                                            _c_double, _c_double, _c_double, _dp, _vp, _vp]),
    "fc3_third_order_sw": (_c_int, [_handle, _dp, _c_int, _dp, _dp, _c_double, _c_double, _dp,
                                    _vp, _vp]),
    "fc3_modes_lj": (_c_int, [_handle, _dp, _c_int, _dp, _c_double, _c_double, _c_double,
                              _c_double, _c_double, _dp, _c_int, _dp, _vp, _vp]),
    "fc3_modes_sw": (_c_int, [_handle, _dp, _c_int, _dp, _dp, _c_double, _c_double, _dp, _c_int,
                              _dp, _vp, _vp]),
    "fc3_third_order_mcc3_lj": (_c_int, [_handle, _dp, _c_int, _dp, _c_double, _c_double,
                                         _c_double, _c_double, _c_double, _dp, _vp, _vp]),
    "fc3_third_order_mcc3_sw": (_c_int, [_handle, _dp, _c_int, _dp, _dp, _c_double, _c_double,
                                         _dp, _vp, _vp]),
    "fc3_mass_weight_third_order": (_c_int, [_handle, _vp, _c_int, _c_int, _dp]),
    "fc3_asr_satisfied": (_c_int, [_handle, _vp, _c_int, _c_int, _c_int, _c_double,
                                   ctypes.POINTER(ctypes.c_longlong),
                                   ctypes.POINTER(ctypes.c_longlong), _dp]),
    "fc3_issymmetric2": (_c_int, [_handle, _vp, _c_int, ctypes.POINTER(ctypes.c_longlong)]),
    "fc3_save_third_order_txt": (_c_int, [_handle, _vp, _c_int, _c_int, ctypes.c_char_p,
                                          ctypes.c_char_p]),
    "fc3_save_second_order_txt": (_c_int, [_handle, _vp, _c_int, _c_int, ctypes.c_char_p,
                                           ctypes.c_char_p]),
    "fc3_format_float": (_c_int, [_c_double, ctypes.c_char_p, _c_int]),
    "fc3_mcc3": (_c_int, [_handle, _vp, _vp, _c_int, _c_int, _vp]),
    "fc3_mcc3_inplace": (_c_int, [_handle, _vp, _vp, _c_int, _c_int]),
    "fc3_mcc3_unique": (_c_int, [_handle, _vp, _vp, _c_int, _c_int, _vp]),
    "fc3_mcc3_blocked": (_c_int, [_handle, _vp, _vp, _c_int, _c_int, _c_int, _fp]),
    "fc3_batch_third_order_mcc3_lj": (_c_int, [_handle, ctypes.POINTER(Snapshot), _c_int, _c_double,
                                               _c_double, _c_double, _c_double, _c_double, _c_int,
                                               _c_int]),
    "fc3_batch_graph_nodes": (_c_int, [_handle, ctypes.POINTER(_c_size)]),
    "fc3_batch_graph_lanes": (_c_int, [_handle, ctypes.POINTER(_c_int)]),
    "fc3_weighted_square_loss": (_c_int, [_handle, _vp, ctypes.c_longlong, ctypes.c_ulonglong,
                                          _vp, ctypes.POINTER(_c_double)]),
    "fc3_autotune_query": (_c_int, [_handle, _c_int, _c_int, ctypes.POINTER(_c_int),
                                    ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_float)]),
    "fc3_autotune_save": (_c_int, [_handle, ctypes.c_char_p]),
    "fc3_autotune_load": (_c_int, [_handle, ctypes.c_char_p]),
    "fc3_third_order_sw_param_backward": (_c_int, [_handle, _dp, _c_int, _dp, _dp, _c_double,
                                                   _c_double, _dp, _vp, _vp, _dp]),
    "fc3_mcc3_backward": (_c_int, [_handle, _vp, _vp, _vp, _c_int, _c_int, _vp, _vp]),
    "fc3_third_order_lj_param_backward": (_c_int, [_handle, _dp, _c_int, _dp, _c_double, _c_double,
                                                   _c_double, _c_double, _c_double, _dp, _vp, _vp,
                                                   _dp]),
}

_lib = None


class FC3Error(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"[fc3 status {status}] {msg}")
        self.status = status


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run forceconstants.jl_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    """Raise the exception the reference raises for the same condition."""
    if status == FC3_OK:
        return
    msg = load().fc3_last_error().decode("utf-8", "replace")
    if status in (FC3_ERR_ARGUMENT, FC3_ERR_CUTOFF, FC3_ERR_BLOCK_SIZE):
        # the reference's @assert failures (third_order_analytical.jl:7,9; mcc.jl:91-92)
        raise AssertionError(msg)
    raise FC3Error(status, msg)
