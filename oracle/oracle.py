"""ctypes front-end of the CPU oracle (oracle/parflex_oracle.c).

TEST INFRASTRUCTURE ONLY: tests/, the smoke check and bench.py's CPU-baseline
leg use this module as the checker.  The product package never imports it.

The functions mirror the reference's Python surface so that parity tests read
like the reference's own:
    flexdtw(C, steps, weights, buffer)          FlexDTW.py:1320-1372
    run_stage1(C, L, steps, weights)            Parflex.py:376-469 (edges only)
    run_stage2(C, L, steps, beta, s1)           Parflex.py:862-1356
    propagate(C, s1)                            Parflex.py:862-912 (before the sync)
    sync_overlap(D_arr, L_arr, edge_lens)       Parflex.py:919-969
    parflex(C, steps, weights, beta, L)         Parflex.py:1510-1547
    cost_matrix(F1, F2)                         Parflex.py:1470
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_parflex.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i8p = ctypes.POINTER(ctypes.c_int8)
_i64 = ctypes.c_int64


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    lib.or_flexdtw_dp.argtypes = [_dp, _i64, _i64, _i64, _i64p, _dp, ctypes.c_int,
                                  _dp, _i32p, _i8p]
    lib.or_flexdtw_dp.restype = None
    lib.or_find_best_endpoint.argtypes = [_dp, _i32p, _i64, _i64, _i64, _dp, _dp,
                                          _i64p, _i64p]
    lib.or_find_best_endpoint.restype = _i64
    lib.or_backtrace.argtypes = [_i8p, _i64, _i64, _i64p, _i64, _i64, _i64p]
    lib.or_backtrace.restype = _i64
    lib.or_stage1.argtypes = [_dp, _i64, _i64, _i64, _i64p, _dp, ctypes.c_int,
                              ctypes.c_int, _dp, _i32p, _i8p, _dp]
    lib.or_stage1.restype = ctypes.c_int
    lib.or_stage2.argtypes = [_dp, _i64, _i64, _i64, _i64p, ctypes.c_int,
                              ctypes.c_double, _dp, _i32p, _i8p, _dp, _dp, _i64p,
                              _dp, _dp, _dp, _dp, _dp, _dp, _dp, _i64p, _i64, _i64p,
                              _dp, _i64p, _i64, _i64p]
    lib.or_stage2.restype = ctypes.c_int
    lib.or_backtrace_and_stitch.argtypes = [_i64, _i64, _i64, _i64p, ctypes.c_int,
                                            _i32p, _i8p, _i64, _i64, _i64, _i64,
                                            _i64p, _i64]
    lib.or_backtrace_and_stitch.restype = _i64
    lib.or_parflex.argtypes = [_dp, _i64, _i64, _i64, _i64p, _dp, ctypes.c_int,
                               ctypes.c_double, ctypes.c_int, _dp, _i64p, _i64, _i64p]
    lib.or_parflex.restype = ctypes.c_int
    lib.or_propagate.argtypes = [_dp, _i64, _i64, _i64, _dp, _i32p, _dp, _dp, _i64p]
    lib.or_propagate.restype = ctypes.c_int
    lib.or_sync_overlap.argtypes = [_dp, _dp, _i64p, _i64, _i64, _i64]
    lib.or_sync_overlap.restype = None
    lib.or_cost_matrix.argtypes = [_dp, _dp, _i64, _i64, _i64, _dp]
    lib.or_cost_matrix.restype = None
    _lib = lib
    return lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _steps_weights(steps, weights):
    s = np.ascontiguousarray(np.asarray(steps, dtype=np.int64).reshape(-1, 2))
    w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64).reshape(-1))
    if s.shape[0] != w.shape[0]:
        raise ValueError("steps and weights differ in length")
    return s, w


def flexdtw_dp(C, steps, weights):
    lib = _load()
    C = np.ascontiguousarray(C, dtype=np.float64)
    R, W = C.shape
    s, w = _steps_weights(steps, weights)
    D = np.empty((R, W), np.float64)
    P = np.empty((R, W), np.int32)
    B = np.empty((R, W), np.int8)
    lib.or_flexdtw_dp(_p(C, _dp), W, R, W, _p(s, _i64p), _p(w, _dp), s.shape[0],
                      _p(D, _dp), _p(P, _i32p), _p(B, _i8p))
    return D, P, B


def find_best_endpoint(D, P, buffer):
    lib = _load()
    D = np.ascontiguousarray(D, dtype=np.float64)
    P = np.ascontiguousarray(P, dtype=np.int32)
    R, W = D.shape
    avg = np.empty(R + W + 2, np.float64)
    bc = ctypes.c_double()
    br = ctypes.c_int64()
    bcc = ctypes.c_int64()
    n = lib.or_find_best_endpoint(_p(D, _dp), _p(P, _i32p), R, W, int(buffer),
                                  _p(avg, _dp), ctypes.byref(bc), ctypes.byref(br),
                                  ctypes.byref(bcc))
    return bc.value, br.value, bcc.value, avg[:n].copy()


def backtrace(B, steps, r, c):
    lib = _load()
    B = np.ascontiguousarray(B, dtype=np.int8)
    R, W = B.shape
    s = np.ascontiguousarray(np.asarray(steps, dtype=np.int64).reshape(-1, 2))
    out = np.empty((R + W + 2, 2), np.int64)
    n = lib.or_backtrace(_p(B, _i8p), R, W, _p(s, _i64p), int(r), int(c), _p(out, _i64p))
    return out[:n].copy()


def flexdtw(C, steps, weights, buffer=1):
    """FlexDTW.flexdtw: (best_cost, wp (2,n), D, P, B, debug)."""
    D, P, B = flexdtw_dp(C, steps, weights)
    R, W = D.shape
    bc, br, bcc, avg = find_best_endpoint(D, P, buffer)
    debug = []
    buffer = int(buffer)
    n_row = max(W - buffer, 0)
    for k, a in enumerate(avg):
        if k < n_row:
            r, c = R - 1, buffer + k
            debug.append((c - W + 1, float(a), r, c))
        else:
            r, c = R - 2 - (k - n_row), W - 1
            debug.append((R - 1 - r, float(a), r, c))
    path = backtrace(B, steps, br, bcc)
    return bc, path.T.copy(), D, P, B, debug


def tile_grid(L1, L2, L):
    hop = L - 1
    return -(-(L1 - 1) // hop), -(-(L2 - 1) // hop)


def tile_bounds(i, j, L1, L2, L):
    hop = L - 1
    s1, s2 = i * hop, j * hop
    return s1, min(s1 + L, L1), s2, min(s2 + L, L2)


def run_stage1(C, L, steps, weights, nthreads=1, keep_B=True, best=False):
    lib = _load()
    C = np.ascontiguousarray(C, dtype=np.float64)
    L1, L2 = C.shape
    n1, n2 = tile_grid(L1, L2, L)
    s, w = _steps_weights(steps, weights)
    eD = np.empty((n1, n2, 2, L), np.float64)
    eP = np.empty((n1, n2, 2, L), np.int32)
    B = np.empty((n1, n2, L, L), np.int8) if keep_B else None
    bst = np.empty((n1, n2), np.float64) if best else None
    rc = lib.or_stage1(_p(C, _dp), L1, L2, L, _p(s, _i64p), _p(w, _dp), s.shape[0],
                       int(nthreads), _p(eD, _dp), _p(eP, _i32p),
                       _p(B, _i8p) if B is not None else None,
                       _p(bst, _dp) if bst is not None else None)
    if rc != 0:
        raise MemoryError("oracle stage 1 failed to allocate")
    return {"L1": L1, "L2": L2, "L": L, "n1": n1, "n2": n2, "edge_D": eD,
            "edge_P": eP, "B": B, "best": bst, "steps": s}


def run_stage2(C, s1, beta):
    lib = _load()
    C = np.ascontiguousarray(C, dtype=np.float64)
    L1, L2, L, n1, n2 = s1["L1"], s1["L2"], s1["L"], s1["n1"], s1["n2"]
    steps = s1["steps"]
    max_segs = max(-(-L2 // L), -(-L1 // L), 1)
    Da = np.empty((n1, n2, 2, L), np.float64)
    La = np.empty((n1, n2, 2, L), np.float64)
    el = np.empty((n1, n2, 2), np.int64)
    tD = np.empty(L2, np.float64)
    tL = np.empty(L2, np.float64)
    rD = np.empty(L1, np.float64)
    rL = np.empty(L1, np.float64)
    sn = np.empty((2, max_segs), np.float64)
    sD = np.empty((2, max_segs), np.float64)
    sL = np.empty((2, max_segs), np.float64)
    si = np.empty((7, 2, max_segs), np.int64)
    bo = np.empty(7, np.int64)
    bn = ctypes.c_double()
    cap = L1 + L2 + 8
    path = np.empty((cap, 2), np.int64)
    plen = ctypes.c_int64()
    rc = lib.or_stage2(_p(C, _dp), L1, L2, L, _p(steps, _i64p), steps.shape[0],
                       float(beta), _p(s1["edge_D"], _dp), _p(s1["edge_P"], _i32p),
                       _p(s1["B"], _i8p), _p(Da, _dp), _p(La, _dp), _p(el, _i64p),
                       _p(tD, _dp), _p(tL, _dp), _p(rD, _dp), _p(rL, _dp),
                       _p(sn, _dp), _p(sD, _dp), _p(sL, _dp), _p(si, _i64p), max_segs,
                       _p(bo, _i64p), ctypes.byref(bn), _p(path, _i64p), cap,
                       ctypes.byref(plen))
    if rc < 0:
        raise MemoryError("oracle stage 2 failed to allocate")
    return {"found": rc == 0, "D_arr": Da, "L_arr": La, "edge_lens": el,
            "best_cost": bn.value, "best_out": bo, "stitched_wp": path[:plen.value].copy(),
            "top_D": tD, "top_L": tL, "right_D": rD, "right_L": rL,
            "seg_norm": sn, "seg_D": sD, "seg_L": sL, "seg_i64": si}


def propagate(C, s1):
    """propagate_tile_edge_costs: (D_arr, L_arr, edge_lens) before the sync."""
    lib = _load()
    C = np.ascontiguousarray(C, dtype=np.float64)
    L1, L2, L, n1, n2 = s1["L1"], s1["L2"], s1["L"], s1["n1"], s1["n2"]
    Da = np.empty((n1, n2, 2, L), np.float64)
    La = np.empty((n1, n2, 2, L), np.float64)
    el = np.empty((n1, n2, 2), np.int64)
    rc = lib.or_propagate(_p(C, _dp), L1, L2, L, _p(s1["edge_D"], _dp), _p(s1["edge_P"], _i32p),
                          _p(Da, _dp), _p(La, _dp), _p(el, _i64p))
    if rc < 0:
        raise MemoryError("oracle propagate failed to allocate")
    return Da, La, el


def sync_overlap(D_arr, L_arr, edge_lens):
    """sync_tile_overlap_edges on copies; returns (D_arr, L_arr)."""
    lib = _load()
    D = np.array(D_arr, dtype=np.float64, order="C", copy=True)
    La = np.array(L_arr, dtype=np.float64, order="C", copy=True)
    el = np.ascontiguousarray(edge_lens, dtype=np.int64)
    n1, n2, _, L = D.shape
    lib.or_sync_overlap(_p(D, _dp), _p(La, _dp), _p(el, _i64p), n1, n2, L)
    return D, La


def backtrace_and_stitch(s1, ci, cj, et, idx):
    lib = _load()
    cap = s1["L1"] + s1["L2"] + 8
    path = np.empty((cap, 2), np.int64)
    n = lib.or_backtrace_and_stitch(s1["L1"], s1["L2"], s1["L"], _p(s1["steps"], _i64p),
                                    s1["steps"].shape[0], _p(s1["edge_P"], _i32p),
                                    _p(s1["B"], _i8p), int(ci), int(cj), int(et), int(idx),
                                    _p(path, _i64p), cap)
    return path[:n].copy()


def parflex(C, steps, weights, beta, L, nthreads=1):
    """Parflex.parflex: (best_cost, wp (2, n))."""
    lib = _load()
    C = np.ascontiguousarray(C, dtype=np.float64)
    L1, L2 = C.shape
    s, w = _steps_weights(steps, weights)
    cap = L1 + L2 + 8
    path = np.empty((cap, 2), np.int64)
    plen = ctypes.c_int64()
    bc = ctypes.c_double()
    rc = lib.or_parflex(_p(C, _dp), L1, L2, int(L), _p(s, _i64p), _p(w, _dp), s.shape[0],
                        float(beta), int(nthreads), ctypes.byref(bc), _p(path, _i64p),
                        cap, ctypes.byref(plen))
    if rc < 0:
        raise MemoryError("oracle parflex failed to allocate")
    if rc == 1:
        raise ValueError("Stage 2: No valid endpoint found on global top/right edges.")
    return bc.value, path[:plen.value].T.copy()


def cost_matrix(F1, F2):
    lib = _load()
    F1 = np.ascontiguousarray(F1, dtype=np.float64)
    F2 = np.ascontiguousarray(F2, dtype=np.float64)
    K, N = F1.shape
    K2, M = F2.shape
    if K != K2:
        raise ValueError("feature dimensions differ")
    C = np.empty((N, M), np.float64)
    lib.or_cost_matrix(_p(F1, _dp), _p(F2, _dp), K, N, M, _p(C, _dp))
    return C
