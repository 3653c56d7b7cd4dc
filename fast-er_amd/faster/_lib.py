"""ctypes binding of libfaster_hip.so (the C ABI declared in include/faster_hip.h).

The library is built in-tree by ``fast-er_amd/build_hip.py`` and loaded from
this directory. There is no fallback: when the library or a ROCm GPU is
missing, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
# FASTER_HIP_LIB names another build of the same library (for timing
# compile-time variants side by side); the default is the in-tree build.
LIB_PATH = os.environ.get("FASTER_HIP_LIB") or os.path.join(_HERE, "libfaster_hip.so")

FH_MODE_LINK = 0
FH_MODE_DEDUP = 1

_lock = threading.Lock()
_lib = None

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int
_f32 = ctypes.c_float
_f64 = ctypes.c_double


class JwStats(ctypes.Structure):
    """Mirror of fh_jw_stats."""

    _fields_ = [
        ("pairs_logical", ctypes.c_int64),
        ("pairs_kernel", ctypes.c_int64),
        ("pairs_pruned", ctypes.c_int64),
        ("pairs_generic", ctypes.c_int64),
        ("alg_lane_ops", ctypes.c_double),
        ("alphabet", ctypes.c_int32),
        ("launches", ctypes.c_int32),
        ("kernel_ms", ctypes.c_double),
        ("method", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("pairs_candidates", ctypes.c_int64),
        ("pass_ms", ctypes.c_double),
        ("eval_ms", ctypes.c_double),
        ("frozen_model_ops", ctypes.c_double),
        ("prefix_skipped_ops", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_SIGS = {
    "fh_last_error": (ctypes.c_char_p, []),
    "fh_version": (_i32, []),
    "fh_probe_valu": (_i32, [_vp, _i32, _i32, ctypes.POINTER(ctypes.c_double), _vp]),
    "fh_jw_set_method": (_i32, [_i32]),
    "fh_jw_get_method": (_i32, []),
    "fh_jw_levels": (_i32, [_vp, _vp, _i64, _vp, _vp, _i64, _i32, _f32, _f32, _f32, _vp, _i64,
                            ctypes.POINTER(_i64), ctypes.POINTER(JwStats), _vp]),
    "fh_expand_link": (_i32, [_vp, _i64, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _i64,
                              ctypes.POINTER(_i64), _vp]),
    "fh_expand_link_sorted": (_i32, [_vp, _i64, _i64, _i64, _vp, _vp, _i64, _vp, _vp, _i64, _vp, _i64,
                                     ctypes.POINTER(_i64), _vp]),
    "fh_expand_dedup": (_i32, [_vp, _i64, _i64, _vp, _vp, _i64, _i32, _vp, _i64,
                               ctypes.POINTER(_i64), _vp]),
    "fh_exact_link": (_i32, [_vp, _i64, _vp, _vp, _i64, _i64, _vp, _i64, ctypes.POINTER(_i64), _vp]),
    "fh_exact_dedup": (_i32, [_vp, _i64, _vp, _vp, _i64, _vp, _i64, ctypes.POINTER(_i64), _vp]),
    "fh_sort_u64": (_i32, [_vp, _i64, _i32, _i32, _vp]),
    "fh_sort_u64_into": (_i32, [_vp, _vp, _i64, _i32, _i32, _vp]),
    "fh_group_strings_multi": (_i32, [_i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fh_group_strings": (_i32, [_vp, _vp, _i64, _vp, _vp, _vp, _vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                                _vp]),
    "fh_group_values": (_i32, [_vp, _i64, _i32, _vp, _vp, _vp, ctypes.POINTER(_i64), _vp]),
    "fh_groups_side_b": (_i32, [_vp, _vp, _i64, _i64, _i64, _vp, _vp, _vp, _vp]),
    "fh_numeric_link": (_i32, [_vp, _i64, _vp, _i64, _f64, _f64, _vp, _i64, ctypes.POINTER(_i64), _vp]),
    "fh_numeric_dedup": (_i32, [_vp, _i64, _f64, _f64, _vp, _i64, ctypes.POINTER(_i64), _vp]),
    "fh_pattern_add_missing": (_i32, [_vp, _vp, _i32, _i64, _vp, _vp, _i64, _i64, _i32, ctypes.POINTER(_i64), _vp,
                                      ctypes.POINTER(_i64), _vp]),
    "fh_split_levels": (_i32, [_vp, _i64, _i32, _vp, ctypes.POINTER(_i64), _vp]),
    "fh_sorted_member_mask": (_i32, [_vp, _i64, _vp, _i64, _i32, _vp, _vp]),
    "fh_pattern_merge": (_i32, [_i32, ctypes.POINTER(_vp), ctypes.POINTER(_i64),
                                ctypes.POINTER(ctypes.c_int32), _vp, _i64, ctypes.POINTER(_i64),
                                ctypes.POINTER(_i64), _vp]),
    "fh_pattern_merge_rowids": (_i32, [_i32, ctypes.POINTER(_vp), ctypes.POINTER(_i64),
                                       ctypes.POINTER(ctypes.c_int32), _i32, ctypes.POINTER(_vp),
                                       ctypes.POINTER(_vp), _i64, _vp, _i64, ctypes.POINTER(_i64),
                                       ctypes.POINTER(_i64), _vp]),
    "fh_pattern_merge_listed": (_i32, [_i32, ctypes.POINTER(_vp), ctypes.POINTER(_i64),
                                       ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                       _i64, _vp, _i64, ctypes.POINTER(_i64), ctypes.POINTER(_i64), _vp]),
    "fh_em_fit": (_i32, [_i32, ctypes.POINTER(ctypes.c_int32), _vp, _vp, _f64, _i32, _vp,
                         ctypes.POINTER(_i32), ctypes.POINTER(_i32), _vp]),
    "fh_em_fit_noalloc": (_i32, [_i32, ctypes.POINTER(ctypes.c_int32), _vp, _vp, _f64, _i32, _vp,
                                 ctypes.POINTER(_i32), ctypes.POINTER(_i32), _vp]),
    "fh_match_probability": (_i32, [_i32, ctypes.POINTER(ctypes.c_int32), _vp, _vp, _vp]),
    "fh_pair_posteriors": (_i32, [_vp, _i32, _vp, _vp, _i64, _vp]),
    "fh_fs_loglik_fwd": (_i32, [_i32, ctypes.POINTER(ctypes.c_int32), _vp, _vp, _vp, _vp, _vp]),
    "fh_fs_loglik_bwd": (_i32, [_i32, ctypes.POINTER(ctypes.c_int32), _vp, _vp, _vp, _vp, _vp]),
    "fh_fs_loglik_hess": (_i32, [_i32, ctypes.POINTER(ctypes.c_int32), _vp, _vp, _vp, _vp]),
}



class BlockedTuning(ctypes.Structure):
    """Mirror of fh_blocked_tuning."""

    _fields_ = [
        ("tile_width", ctypes.c_int32),
        ("tiles_per_item", ctypes.c_int32),
        ("waves_per_item", ctypes.c_int32),
        ("wgs_per_cu", ctypes.c_int32),
        ("hash_slots", ctypes.c_int64),
    ]


_SIGS.update({
    "fh_blocked_plan_create": (_i32, [_i32, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _i64, ctypes.POINTER(_vp),
                                      ctypes.POINTER(_vp), _i64, _vp, _vp, ctypes.c_int32, _i32,
                                      ctypes.POINTER(_vp), ctypes.POINTER(_vp), _f32, _f32, _f32,
                                      ctypes.POINTER(BlockedTuning), ctypes.POINTER(_vp), _vp]),
    "fh_blocked_run": (_i32, [_vp, _i32, _vp]),
    "fh_blocked_run_per_block": (_i32, [_vp, _vp]),
    "fh_blocked_counts": (_i32, [_vp, _vp, ctypes.POINTER(_i64), ctypes.POINTER(_i32), ctypes.POINTER(_f64)]),
    "fh_blocked_info": (_i32, [_vp, ctypes.POINTER(_f64)]),
    "fh_blocked_plan_destroy": (_i32, [_vp]),
})

EXPORTED = tuple(_SIGS)


def load(path: str = LIB_PATH, require_gpu: bool = True):
    """Load the library (once). With ``require_gpu`` the call fails unless a
    ROCm device is visible to torch."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(path):
                raise RuntimeError(
                    f"libfaster_hip.so not found at {path}; build it with "
                    "`python fast-er_amd/build_hip.py`")
            L = ctypes.CDLL(path)
            for name, (res, args) in _SIGS.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    if require_gpu and not torch.cuda.is_available():
        raise RuntimeError("faster (MI355X build) needs a ROCm GPU; none is visible")
    return _lib


def check(status: int, what: str) -> None:
    if status != 0:
        msg = _lib.fh_last_error().decode(errors="replace") if _lib is not None else ""
        raise RuntimeError(f"{what} failed (status {status}): {msg}")


def stream():
    return _vp(torch.cuda.current_stream().cuda_stream)


def ptr(t: torch.Tensor | None):
    if t is None:
        return _vp(0)
    return _vp(t.data_ptr())


def device() -> torch.device:
    return torch.device("cuda", torch.cuda.current_device())


FH_JW_METHOD_TILES = 0
FH_JW_METHOD_PREFIX = 2
_METHODS = {"tiles": FH_JW_METHOD_TILES, "prefix": FH_JW_METHOD_PREFIX}


def set_jw_method(name: str) -> None:
    """Selects the all-pairs comparison algorithm of every later call
    (process-wide): "prefix" (the default: bit-parallel greedy match with
    each wave walking its sorted string-1 values against a staged string-2
    tile and reusing shared prefixes) or "tiles" (the same match with one
    string-1 value per lane). Both give identical results."""
    if name not in _METHODS:
        raise ValueError(f"unknown comparison method {name!r}; expected one of {sorted(_METHODS)}")
    check(load().fh_jw_set_method(_METHODS[name]), "fh_jw_set_method")


def jw_method() -> str:
    m = load().fh_jw_get_method()
    return {v: k for k, v in _METHODS.items()}[m]
