"""ctypes binding of ``libcsg.so`` (declared in ``include/csg.h``).

Loading fails loudly when the library has not been built: there is no CPU
fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcsg.so")

CSG_OK = 0
CSG_ERR_ARG = 1
CSG_ERR_HIP = 2
CSG_ERR_NONFINITE = 3
CSG_ERR_UNSUPPORTED = 4
CSG_ERR_STATE = 5

CSG_FP64 = 0
CSG_MIXED = 1

CSG_NOALLOC_MAX_N = 1024
CSG_NOALLOC_SCRATCH_BYTES = 26009344

EXPORTS = (
    "csg_default_params", "csg_create", "csg_destroy", "csg_set_fields", "csg_run",
    "csg_step", "csg_get_field", "csg_nonfinite_step", "csg_chemical_potential",
    "csg_transport_rate", "csg_energy", "csg_geometry", "csg_bulk_energy",
    "csg_bulk_derivative", "csg_initial_fields", "csg_julia_rand", "csg_time_passes",
    "csg_device_bytes", "csg_last_error", "csg_adjoint_forward", "csg_adjoint_backward",
    "csg_batch_create", "csg_batch_destroy", "csg_batch_set_fields", "csg_batch_run",
    "csg_batch_get_field", "csg_batch_nonfinite_step", "csg_batch_device_bytes",
    "csg_schedule_info", "csg_get_schedule", "csg_set_schedule", "csg_autotune",
    "csg_batch_get_schedule", "csg_batch_set_schedule", "csg_batch_autotune", "csg_set_params",
    "csg_copy_bandwidth", "csg_step_noalloc", "csg_batch_set_persistent",
)


class CsgParams(ctypes.Structure):
    _fields_ = [(name, ctypes.c_double) for name in
                ("A", "B", "chi", "P", "ca", "cb", "cc", "kappa", "mobility", "Q")]


class CsgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


_lib = None
_D = ctypes.POINTER(ctypes.c_double)
_VP = ctypes.c_void_p


def lib():
    """The loaded library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run build.py (HIP path required, no fallback)")
    L = ctypes.CDLL(LIB_PATH)
    i32, i64, dbl = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    sig = {
        "csg_default_params": (None, [ctypes.POINTER(CsgParams)]),
        "csg_create": (ctypes.c_int, [ctypes.POINTER(_VP), i32, dbl, ctypes.POINTER(CsgParams), i32, i32]),
        "csg_destroy": (ctypes.c_int, [_VP]),
        "csg_set_fields": (ctypes.c_int, [_VP, _D, _D]),
        "csg_run": (ctypes.c_int, [_VP, dbl, i64, i64, _D]),
        "csg_step": (ctypes.c_int, [_VP, dbl, i64]),
        "csg_get_field": (ctypes.c_int, [_VP, _D]),
        "csg_nonfinite_step": (i64, [_VP]),
        "csg_chemical_potential": (ctypes.c_int, [_VP, _D, _D]),
        "csg_transport_rate": (ctypes.c_int, [_VP, _D, _D, _D, _D]),
        "csg_energy": (ctypes.c_int, [_VP, _D, _D]),
        "csg_geometry": (ctypes.c_int, [_VP, _D, _D, _D]),
        "csg_bulk_energy": (ctypes.c_int, [_VP, _D, _D]),
        "csg_bulk_derivative": (ctypes.c_int, [_VP, _D, _D]),
        "csg_initial_fields": (ctypes.c_int, [i32, dbl, dbl, dbl, i64, _D, _D]),
        "csg_julia_rand": (ctypes.c_int, [i64, i64, _D]),
        "csg_time_passes": (ctypes.c_int, [_VP, dbl, i64, _D]),
        "csg_device_bytes": (i64, [_VP]),
        "csg_adjoint_forward": (ctypes.c_int, [_VP, dbl, i64, i64]),
        "csg_adjoint_backward": (ctypes.c_int, [_VP, _D, _D, _D]),
        "csg_last_error": (ctypes.c_char_p, []),
        "csg_batch_create": (ctypes.c_int, [ctypes.POINTER(_VP), i32, ctypes.POINTER(i32), _D,
                                            ctypes.POINTER(CsgParams), i32, i32]),
        "csg_batch_destroy": (ctypes.c_int, [_VP]),
        "csg_batch_set_fields": (ctypes.c_int, [_VP, i32, _D, _D]),
        "csg_batch_run": (ctypes.c_int, [_VP, dbl, i64, i64, ctypes.POINTER(_D)]),
        "csg_batch_get_field": (ctypes.c_int, [_VP, i32, _D]),
        "csg_batch_nonfinite_step": (i64, [_VP, i32]),
        "csg_batch_device_bytes": (i64, [_VP]),
        "csg_schedule_info": (ctypes.c_int, [_VP, i32, ctypes.POINTER(i32), ctypes.POINTER(i32)]),
        "csg_get_schedule": (ctypes.c_int, [_VP, ctypes.POINTER(i32)]),
        "csg_set_schedule": (ctypes.c_int, [_VP, ctypes.POINTER(i32)]),
        "csg_autotune": (ctypes.c_int, [_VP, dbl, i32, ctypes.POINTER(i32), _D]),
        "csg_batch_get_schedule": (ctypes.c_int, [_VP, i32, ctypes.POINTER(i32)]),
        "csg_batch_set_schedule": (ctypes.c_int, [_VP, i32, ctypes.POINTER(i32)]),
        "csg_batch_autotune": (ctypes.c_int, [_VP, dbl, i32]),
        "csg_set_params": (ctypes.c_int, [_VP, ctypes.POINTER(CsgParams)]),
        "csg_copy_bandwidth": (ctypes.c_int, [i32, i64, i32, _D]),
        "csg_batch_set_persistent": (ctypes.c_int, [_VP, i32]),
        "csg_step_noalloc": (ctypes.c_int, [i32, dbl, ctypes.POINTER(CsgParams), i32, i32, dbl, i64, _VP, _VP, _VP]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def ptr(a: np.ndarray):
    """Pointer to a float64 Fortran- or C-contiguous array (caller keeps it alive)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.flags.c_contiguous)
    return a.ctypes.data_as(_D)


def check(rc: int):
    if rc != CSG_OK:
        msg = lib().csg_last_error().decode(errors="replace")
        raise CsgError(rc, msg or f"csg error {rc}")
