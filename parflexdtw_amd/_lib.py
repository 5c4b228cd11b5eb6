"""ctypes binding of libparflex_hip.so (the C ABI declared in include/parflex_hip.h).

The product path has no CPU fallback: if the library is missing or cannot be
loaded, every call raises RuntimeError naming the problem.
"""
from __future__ import annotations

import ctypes
import os
import threading

__all__ = ["lib", "check", "LIB_PATH", "PF_F64", "PF_F32", "PF_BF16", "ParflexError"]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libparflex_hip.so")

PF_F64, PF_F32, PF_BF16 = 0, 1, 2
PF_S2_PROPAGATE, PF_S2_SYNC, PF_S2_SCAN, PF_S2_STITCH = 1, 2, 4, 8
PF_S2_ALL = PF_S2_PROPAGATE | PF_S2_SYNC | PF_S2_SCAN | PF_S2_STITCH
PF_EINVAL, PF_EHIP, PF_ENOENDPOINT = -1, -2, -3

_lock = threading.Lock()
_lib = None

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double
_flt = ctypes.c_float

# name -> (restype, argtypes)
_SIGNATURES = {
    "pf_last_error": (ctypes.c_char_p, []),
    "pf_version": (_int, []),
    "pf_valu_probe": (_int, [_vp, _int, _int, _vp]),
    "pf_check_steps": (_int, [_vp, _vp, _int]),
    "pf_stage1": (_int, [_int, _vp, _vp, _i64, _i64, _vp, _vp, _int, _int, _int, _int, _vp, _vp,
                         _vp, _vp, _vp, _vp, _vp]),
    "pf_stage1_noalloc": (_int, [_int, _vp, _vp, _i64, _i64, _vp, _vp, _int, _int, _int, _int,
                                 _vp, _vp, _vp, _vp, _vp]),
    "pf_stage1_workspace": (_int, [_int, _i64, _i64, _int, _vp, _vp]),
    "pf_stage1_features": (_int, [_vp, _vp, _int, _vp, _i64, _i64, _vp, _vp, _int, _int, _int,
                                  _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pf_set_wait_limit": (_int, [_i64, _vp]),
    "pf_device_status": (_int, [_vp]),
    "pf_features_bf16": (_int, [_vp, _i64, _i64, _vp, _vp]),
    "pf_tile_full": (_int, [_vp, _i64, _i64, _i64, _i64, _i64, _vp, _vp, _int, _vp, _vp, _vp,
                            _vp]),
    "pf_tile_endpoint": (_int, [_vp, _vp, _i64, _i64, _i64, _vp, _vp, _vp]),
    "pf_tile_backtrace": (_int, [_vp, _i64, _i64, _vp, _int, _vp, _vp, _vp, _vp]),
    "pf_stage2": (_int, [_vp, _vp, _i64, _vp, _i64, _i64, _i64, _i64, _i64, _i64, _vp,
                         _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                         _vp, _vp, _vp, _vp, _vp, _vp]),
    "pf_stage2_phases": (_int, [_int, _vp, _vp, _i64, _vp, _i64, _i64, _i64, _i64, _i64, _i64,
                                _vp, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pf_stage2_propagate_listed": (_int, [_vp, _vp, _i64, _vp, _i64, _i64, _vp, _vp, _i64, _vp,
                                          _vp, _vp, _vp, _vp, _vp, _vp]),
    "pf_stage2_propagate_dag": (_int, [_vp, _vp, _i64, _vp, _i64, _i64, _vp, _vp, _i64, _vp,
                                       _vp, _vp, _vp, _vp, _vp, _vp]),
    "pf_stitch": (_int, [_vp, _i64, _vp, _vp, _i64, _vp, _int, _vp, _vp, _vp, _vp, _i64, _vp,
                         _vp, _vp]),
    "pf_cost": (_int, [_int, _vp, _vp, _i64, _i64, _i64, _vp, _i64, _vp, _vp]),
    "pf_cost_workspace": (_int, [_int, _i64, _i64, _i64, _vp]),
    "pf_cost_tile": (_int, [_int, _vp, _vp]),
    "pf_cost_batch": (_int, [_int, _vp, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _vp, _vp]),
    "pf_pool_features": (_int, [_vp, _i64, _i64, _int, _vp, _vp]),
    "pf_pack_paths": (_int, [_vp, _i64, _i64, _vp, _vp, _vp, _vp]),
    "pf_pool_features_batch": (_int, [_vp, _i64, _i64, _vp, _i64, _int, _vp, _i64, _vp]),
    "pf_pool_features_bf16_batch": (_int, [_vp, _i64, _i64, _vp, _i64, _int, _vp, _i64, _vp,
                                           _vp]),
    "pf_input_stats_batch": (_int, [_vp, _i64, _vp, _i64, _i64, _vp, _i64, _vp, _vp]),
    "pf_conv3x3_first_pooled": (_int, [_vp, _i64, _i64, _i64, _vp, _i64, _i64, _i64, _i64, _vp,
                                       _vp, _vp, _vp, _vp]),
    "pf_pooled_cost_workspace": (_int, [_i64, _vp]),
    "pf_pooled_cost": (_int, [_vp, _i64, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp]),
    "pf_conv3x3": (_int, [_int, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp]),
    "pf_conv3x3_wgrad_workspace": (_int, [_i64, _i64, _vp]),
    "pf_conv3x3_wgrad": (_int, [_vp, _vp, _i64, _i64, _vp, _vp, _vp]),
    "pf_bce": (_int, [_vp, _vp, _i64, _flt, _flt, _vp, _vp, _vp]),
    "pf_adam": (_int, [_vp, _vp, _vp, _vp, _vp, _i64, _flt, _flt, _flt, _flt, _i64, _vp]),
    "pf_pack_weights": (_int, [_vp, _vp, _vp, _vp]),
    "pf_tile_scores": (_int, [_vp, _i64, _i64, _int, _i64, _i64, _i64, _int, _vp, _vp]),
    "pf_tile_net_geometry": (_int, [_vp, _vp]),
    "pf_tile_net": (_int, [_vp, _i64, _vp, _i64, _i64, _vp, _vp, _i64, _vp, _i64, _vp, _vp, _vp,
                           _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pf_tile_scores_batch": (_int, [_vp, _vp, _vp, _i64, _int, _i64, _int, _vp, _vp]),
    "pf_selection_check": (_int, [_vp, _i64, _i64, _vp, _vp, _vp, _i64, _i64, _vp, _vp]),
    "pf_select_tiles": (_int, [_vp, _dbl, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _vp,
                               _vp, _vp, _vp, _vp, _vp]),
    "pf_eval_paths": (_int, [_vp, _i64, _vp, _vp, _i64, _vp, _i64, _vp, _vp, _dbl, _vp, _i64,
                             _vp, _vp, _vp]),
    "pf_fill_f64": (_int, [_vp, _dbl, _i64, _vp]),
    "pf_fill_i64": (_int, [_vp, _i64, _i64, _vp]),
}

EXPORTS = tuple(_SIGNATURES)


class ParflexError(RuntimeError):
    """A C-ABI call returned an error status."""


def lib():
    """Load (once) and return the HIP library; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP extension first "
                "(python -c 'import __graft_entry__ as g; g.build()' or make -C "
                "parflexdtw_amd/csrc)")
        try:
            h = ctypes.CDLL(LIB_PATH)
        except OSError as exc:
            raise RuntimeError(f"cannot load {LIB_PATH}: {exc}") from exc
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
        return _lib


def last_error():
    msg = lib().pf_last_error()
    return msg.decode() if msg else ""


def check(rc, what):
    """Map a pf_status to a Python exception (ValueError for bad arguments,
    ParflexError for runtime failures)."""
    if rc == 0:
        return
    msg = last_error()
    if rc == PF_EINVAL:
        raise ValueError(f"{what}: {msg}")
    raise ParflexError(f"{what} failed ({rc}): {msg}")
