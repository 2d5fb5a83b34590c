"""ctypes binding of the C ABI in include/mpf.h (libmpf_hip.so, built in-tree).

There is no fallback: if the library is missing or fails to load, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libmpf_hip.so"

MPF_OK = 0
MPF_ERR_ARG, MPF_ERR_UNSUPPORTED, MPF_ERR_HIP, MPF_ERR_NOMEM = 1, 2, 3, 4
MPF_SLOT_OVERFLOW, MPF_EMPTY_NODE, MPF_ERR_STATE = 5, 6, 7
STATUS_NAMES = {
    0: "MPF_OK", 1: "MPF_ERR_ARG", 2: "MPF_ERR_UNSUPPORTED", 3: "MPF_ERR_HIP", 4: "MPF_ERR_NOMEM",
    5: "MPF_SLOT_OVERFLOW", 6: "MPF_EMPTY_NODE", 7: "MPF_ERR_STATE",
}
MPF_FP64, MPF_FP32 = 0, 1
MPF_TJ_MODEL, MPF_TJ_VALUE, MPF_TJ_NONE = 0, 1, 2
MPF_ENERGY_ISO, MPF_ENERGY_CUBIC = 0, 1
MPF_DG_WEIGHTS_PHASEFIELDS, MPF_DG_WEIGHTS_RANGE, MPF_DG_WEIGHTS_COUNTER = 0, 1, 2
MPF_IP_NONE, MPF_IP_CURRENT, MPF_IP_PREVIOUS = 0, 1, 2

# Every symbol the header declares (checked by the CPU tests without calling into the GPU).
EXPORTED = [
    "mpf_create", "mpf_set_grains", "mpf_upload", "mpf_finalize_initialization", "mpf_step",
    "mpf_step_async", "mpf_synchronize", "mpf_total_entries", "mpf_download", "mpf_get_grains",
    "mpf_get_grain_records",
    "mpf_max_time_step", "mpf_interface_energy", "mpf_phase_fractions", "mpf_set_step_terms", "mpf_limiter_warnings",
    "mpf_overflow_cell", "mpf_device_state",
    "mpf_last_timing", "mpf_enable_timing", "mpf_step_stats", "mpf_copy_bandwidth", "mpf_set_pair_params",
    "mpf_state_bytes", "mpf_save_state", "mpf_restore_state", "mpf_adjoint_step", "mpf_autotune",
    "mpf_set_launch_shape", "mpf_set_tile_shape", "mpf_get_tile_shape", "mpf_ensemble_create", "mpf_ensemble_step", "mpf_ensemble_destroy",
    "mpf_ensemble_error_message", "mpf_step_into", "mpf_error_message", "mpf_destroy",
]


class MpfGrid(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("dx", C.c_double), ("iwidth", C.c_double),
        ("bc", C.c_int32 * 6),
        ("precision", C.c_int32), ("slots", C.c_int32),
    ]


class MpfParams(C.Structure):
    _fields_ = [
        ("nphases", C.c_int32),
        ("sigma", C.POINTER(C.c_double)), ("mu", C.POINTER(C.c_double)),
        ("energy_model", C.POINTER(C.c_int32)), ("epsilon", C.POINTER(C.c_double)),
        ("tj_model", C.c_int32), ("tj_factor", C.c_double), ("tj_energy", C.c_double),
        ("consider_nucleus_volume", C.c_int32), ("nucleus_volume_factor", C.c_double),
        ("dg_mode", C.POINTER(C.c_int32)), ("dg_value", C.POINTER(C.c_double)),
        ("dg_limit", C.c_double), ("regularization_factor", C.c_double),
        ("dg_average", C.c_int32), ("dg_range", C.c_int32), ("dg_threshold", C.c_double),
        ("dg_weights", C.c_int32),
        ("regularization", C.c_int32), ("curvature_factor", C.c_double),
        ("reg_average", C.c_int32), ("reg_range", C.c_int32), ("reg_threshold", C.c_double),
    ]


class MpfError(RuntimeError):
    def __init__(self, status: int, where: str, message: str = ""):
        self.status = status
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)} {message}".strip())


_lib = None


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Loads libmpf_hip.so once; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is not built; run __graft_entry__.build() or python openphase_amd/build.py hip")
    lib = C.CDLL(str(p))
    vp = C.c_void_p
    i32p = C.POINTER(C.c_int32)
    f64p = C.POINTER(C.c_double)
    sig = {
        "mpf_create": ([C.POINTER(MpfGrid), C.POINTER(MpfParams), C.POINTER(vp)], C.c_int),
        "mpf_set_grains": ([vp, C.c_int32, i32p, f64p, C.POINTER(C.c_uint8)], C.c_int),
        "mpf_upload": ([vp, i32p, i32p, f64p], C.c_int),
        "mpf_finalize_initialization": ([vp], C.c_int),
        "mpf_step": ([vp, C.c_double, C.c_int32], C.c_int),
        "mpf_step_async": ([vp, C.c_double, C.c_int32], C.c_int),
        "mpf_synchronize": ([vp], C.c_int),
        "mpf_total_entries": ([vp, C.POINTER(C.c_int64)], C.c_int),
        "mpf_download": ([vp, i32p, i32p, f64p, C.POINTER(C.c_int8)], C.c_int),
        "mpf_get_grains": ([vp, f64p, f64p], C.c_int),
        "mpf_get_grain_records": ([vp, f64p, f64p, C.POINTER(C.c_uint8), C.POINTER(C.c_int32)], C.c_int),
        "mpf_max_time_step": ([vp, f64p], C.c_int),
        "mpf_interface_energy": ([vp, C.c_int32, f64p], C.c_int),
        "mpf_set_step_terms": ([vp, C.c_int32, C.c_int32], C.c_int),
        "mpf_phase_fractions": ([vp, f64p, C.c_int32], C.c_int),
        "mpf_limiter_warnings": ([vp, C.POINTER(C.c_int64)], C.c_int),
        "mpf_overflow_cell": ([vp, C.POINTER(C.c_int64)], C.c_int),
        "mpf_device_state": ([vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)], C.c_int),
        "mpf_last_timing": ([vp, f64p, f64p], C.c_int),
        "mpf_enable_timing": ([vp, C.c_int32], C.c_int),
        "mpf_copy_bandwidth": ([vp, C.c_int64, C.c_int32, f64p], C.c_int),
        "mpf_step_stats": ([vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
        "mpf_set_pair_params": ([vp, f64p, f64p, f64p], C.c_int),
        "mpf_state_bytes": ([vp, C.POINTER(C.c_int64)], C.c_int),
        "mpf_save_state": ([vp, vp], C.c_int),
        "mpf_restore_state": ([vp, vp], C.c_int),
        "mpf_adjoint_step": ([vp, C.c_double, vp, vp, vp, vp, vp], C.c_int),
        "mpf_autotune": ([vp, C.c_double, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
        "mpf_set_launch_shape": ([vp, C.c_int32, C.c_int32], C.c_int),
        "mpf_set_tile_shape": ([vp, C.c_int32, C.c_int32, C.c_int32], C.c_int),
        "mpf_get_tile_shape": ([vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
        "mpf_ensemble_create": ([C.POINTER(vp), C.c_int32, C.POINTER(vp)], C.c_int),
        "mpf_ensemble_step": ([vp, f64p, C.c_int32], C.c_int),
        "mpf_ensemble_destroy": ([vp], None),
        "mpf_ensemble_error_message": ([vp], C.c_char_p),
        "mpf_step_into": ([vp, C.c_double, vp, vp, vp, vp, vp, vp], C.c_int),
        "mpf_error_message": ([vp], C.c_char_p),
        "mpf_destroy": ([vp], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def ptr(a: np.ndarray | None, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))
