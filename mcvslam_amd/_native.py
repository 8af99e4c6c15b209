"""ctypes binding of the C ABI in include/mcvslam_gpu.h.

The HIP library is the only implementation of the hot path: if it has not
been built, or no GPU is visible, every entry point raises instead of
falling back to anything else.
"""
from __future__ import annotations

import ctypes
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libmcvslam_gpu.so"

P = ctypes.c_void_p
I = ctypes.c_int
F = ctypes.c_float
I64 = ctypes.c_int64

MCV_OK = 0
MCV_ERR_INVALID = -1
MCV_ERR_HIP = -2
MCV_ERR_CAPACITY = -3
MCV_ERR_EMPTY = -4
MCV_ERR_STATE = -5
MCV_FLAG_GRAPH = 1
MCV_FLAG_CHECK = 2


class KeyPoint(ctypes.Structure):
    _fields_ = [("x", F), ("y", F), ("size", F), ("angle", F), ("response", F), ("octave", ctypes.c_int32)]


# name -> (restype, argtypes)
SIGNATURES = {
    "mcv_create": (I, [ctypes.POINTER(P), I]),
    "mcv_destroy": (None, [P]),
    "mcv_last_error": (ctypes.c_char_p, [P]),
    "mcv_set_stream": (I, [P, P]),
    "mcv_use_own_stream": (I, [P]),
    "mcv_synchronize": (I, [P]),
    "mcv_copy_to_host": (I, [P, P, P, ctypes.c_uint64]),
    "mcv_orb_configure": (I, [P, I, F, I, I, I]),
    "mcv_orb_tables": (I, [P, P, P, P, P, P]),
    "mcv_orb_capacity": (I, [P, I, I, I, P]),
    "mcv_orb_extract": (I, [P, P, I, I, I, P, I, P, P, I, P]),
    "mcv_orb_pyramid": (I, [P, I, P, I, P, P]),
    "mcv_orb_extract_batch_dev": (I, [P, I, P, P, P, P, P, P, P, P, I, P, I]),
    "mcv_orb_batch_check": (I, [P]),
    "mcv_orb_batch_pyramid_dev": (I, [P, I, I, I, P, P, P, P]),
    "mcv_rig_pyramid_dev": (I, [P, I, I, I, P, P, P, P]),
    "mcv_match_knn2": (I, [P, P, I, P, I, P]),
    "mcv_match_set_variant": (I, [P, I]),
    "mcv_distinctive_descriptors": (I, [P, P, P, I, P, P]),
    "mcv_match_filter_orientation": (I, [P, P, I, P, I, P, I, P, P]),
    "mcv_rig_autotune": (I, [P, P, I, P, I]),
    "mcv_rig_set_variants": (I, [P, I, I, I, I]),
    "mcv_rig_get_variants": (I, [P, P, P, P, P]),
    "mcv_match_knn2_dev": (I, [P, P, I, P, I, P]),
    "mcv_match_filter": (I, [P, P, I, F, I, P, P]),
    "mcv_match_filter_noalloc_dev": (I, [P, P, I, F, I, P, P]),
    "mcv_rig_create": (I, [P, I, P, P, P, F, I, P]),
    "mcv_rig_destroy": (None, [P]),
    "mcv_rig_run": (I, [P, P, I]),
    "mcv_rig_outputs": (I, [P, P, P, P, P, P, P, P]),
    "mcv_rig_run_timed": (I, [P, P, P, P, I, P]),
    "mcv_rig_check": (I, [P]),
    "mcv_rig_octree_stats": (I, [P, P, P, P]),
    "mcv_rig_stereo": (I, [P, F, F]),
    "mcv_rig_stereo_outputs": (I, [P, P, P, P]),
    "mcv_window_search": (I, [P, P, P, P, I, P, P, P, P, I, F, F, I, P, P, I, I, I, P]),
    "mcv_window_search_dev": (I, [P, I, P, P, P, I, P, P, P, P, I, F, F, I, P, P, P, P]),
    "mcv_window_search_uv_dev": (I, [P, P, P, I, F, I, I, F, I, P, P, P, I, I, I, P]),
    "mcv_bow_vocab_create": (I, [P, I, P, P, P, P, P, I, I, I, P]),
    "mcv_bow_vocab_destroy": (None, [P]),
    "mcv_bow_transform_dev": (I, [P, P, I, I, P, P, P]),
    "mcv_bow_transform": (I, [P, P, I, I, P, P, P, P, P, P, P, P, P, P]),
    "mcv_bow_match": (I, [P, P, I, I, P, P, P, P, I, I, P, P, P, P, P]),
    "mcv_ba_create": (I, [P, I, I, I, I, P]),
    "mcv_ba_destroy": (None, [P]),
    "mcv_ba_set_observations": (I, [P, P, P, P, P, P]),
    "mcv_ba_set_state": (I, [P, P, P, P]),
    "mcv_ba_linearize": (I, [P, I]),
    "mcv_ba_set_variant": (I, [P, I]),
    "mcv_ba_autotune": (I, [P, I, P, I]),
    "mcv_ba_linearize_plain": (I, [P, I]),
    "mcv_ba_set_stereo": (I, [P, I, P, P, P, P, P]),
    "mcv_ba_linearize_stereo": (I, [P, I]),
    "mcv_ba_read_stereo_planes": (I, [P, I, I, P]),
    "mcv_ba_outliers": (I, [P, F, F, P, P]),
    "mcv_ba_generation": (I, [P, P]),
    "mcv_ba_planes": (I, [P, P, P]),
    "mcv_ba_read_planes": (I, [P, I, I, P]),
    "mcv_ba_copy_planes_dev": (I, [P, I, I, P]),
    "mcv_ba_adjoint_dev": (I, [P, P, P, P]),
    "mcv_ba_set_fixed": (I, [P, P]),
    "mcv_ba_schur": (I, [P, F]),
    "mcv_ba_read_schur": (I, [P, P, P, P]),
    "mcv_solve_llt": (I, [I, P, P, P]),
    "mcv_ba_set_point_fixed": (I, [P, P]),
    "mcv_ba_set_active": (I, [P, P, P]),
    "mcv_ba_set_robust": (I, [P, I]),
    "mcv_ba_robust_chi2": (I, [P, P]),
    "mcv_ba_optimize": (I, [P, I, ctypes.c_double, I, P, P]),
    "mcv_ba_get_state": (I, [P, P, P]),
}

_lib = None
_lock = threading.Lock()


class MCVError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib() -> ctypes.CDLL:
    """Load the HIP library (raises if it is missing)."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build it with `python -m mcvslam_amd.build` "
                    "(there is no CPU implementation of the hot path)")
            l = ctypes.CDLL(str(LIB_PATH))
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(l, name)
                fn.restype = res
                fn.argtypes = args
            _lib = l
        return _lib


def check(ctx, code: int):
    if code != MCV_OK:
        msg = lib().mcv_last_error(ctx).decode() if ctx else "error"
        raise MCVError(code, msg)
    return code
