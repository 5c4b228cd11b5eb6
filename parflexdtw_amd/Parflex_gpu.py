"""MI355X drop-in for the reference's ``Parflex_gpu`` module.

Import it in place of the reference module:

    from parflexdtw_amd import Parflex_gpu as Parflex_gpu
    best_cost, wp = Parflex_gpu.parflex(C, steps, weights, beta, L)

Entry points (same names, arguments, defaults and return values as
Parflex_gpu.py / Parflex.py):

    parflex(C, steps, weights, beta, L=None, profile_dir='Profiling_results',
            backtrace_segments=False, warn_slow_chunks=False,
            slow_chunk_seconds=1.0, use_gpu=None) -> (best_cost, wp (2, n))
                                                   Parflex_gpu.py:656-854
    tiled_stage1_from_features(F1, F2, steps=None, weights=None, beta=0.1,
            L=None, use_gpu=None) -> (C or None, tiled_result)
                                                   Parflex_gpu.py:610-653
    run_stage2_from_tiled(tiled_result, C=None, beta=0.1, show_fig=False,
            top_k=1, profile_dir=None, backtrace_segments=False) -> dict
                                                   Parflex_gpu.py:857-910
    run_flexdtw_on_tiles(C, L, steps=None, weights=None, buffer=1, ...,
            use_gpu=None, C_dev=None) -> (chunks_dict, L, n1, n2)
                                                   Parflex_gpu.py:342-607
    propagate_tile_edge_costs(chunks_dict, L, num_chunks_1, num_chunks_2,
            buffer_param=0.1, profile_dir='Profiling_results', C_dev=None)
            -> (D_arr, L_arr, edge_lens)           Parflex_gpu.py:247-339
    sync_tile_overlap_edges(D_arr, L_arr, edge_lens, n_chunks_1, n_chunks_2)
            -> (D_arr, L_arr), in place            Parflex.py:919-969
    stage2_scan_and_stitch(tiled_result, chunks_dict, D_arr, L_arr, edge_lens,
            L1, L2, L_block, buffer_stage2=200, top_k=1,
            profile_dir='Profiling_results', backtrace_segments=False) -> dict
                                                   Parflex.py:1229-1356
    backtrace_and_stitch(tiled_result, chunks_dict, start_i, start_j,
            start_edge, start_idx) -> [(row, col), ...]
                                                   Parflex.py:1032-1103

The four stage-2 functions take the chunks_dict of run_flexdtw_on_tiles (the
reference's per-tile D, S, B host arrays, from this package or from the
reference) or the device-backed one of tiled_stage1_from_features; they run
the same device kernels as parflex(), one phase per call.

Everything runs on the GPU: stage 1 is one launch over all tiles, stage 2
(propagation, sync, scan, stitch, backtrace) runs on the device, and
parflex() copies back only the best cost and the path.  There is no CPU
implementation in this package: use_gpu=False, or PARFLEX_DISABLE_GPU set with
use_gpu=None, raises RuntimeError (the reference would fall back to its Numba
code, which this package does not ship).
"""
from __future__ import annotations

import csv
import math
import os
import time
import types

import numpy as np

from . import engine
from ._lib import PF_S2_PROPAGATE, PF_S2_SCAN, PF_S2_STITCH, PF_S2_SYNC
from ._plan import BatchPlan, buffer_global, tile_bounds, tile_grid
from .engine import NO_ENDPOINT_MSG, StepWeights
from .gpu_flexdtw import (cost_matrix_from_features_gpu, cost_matrix_to_gpu,
                          cost_matrix_to_gpu_f32, flexdtw_chunk_from_global_C,
                          gather_c_at_global_indices, is_gpu_available)

__all__ = ["parflex", "tiled_stage1_from_features", "run_stage2_from_tiled",
           "run_flexdtw_on_tiles", "propagate_tile_edge_costs", "sync_tile_overlap_edges",
           "stage2_scan_and_stitch", "backtrace_and_stitch", "build_tiled_metadata",
           "edge_slot_to_local", "local_to_global_cell", "global_cell_to_prev_tile_edge", "DEFAULT_CHUNK_LENGTH",
           "cost_matrix_from_features_gpu", "cost_matrix_to_gpu_f32",
           "flexdtw_chunk_from_global_C", "gather_c_at_global_indices", "is_gpu_available"]

DEFAULT_CHUNK_LENGTH = 4000  # Parflex.py:126 (the effective value)
_DEFAULT_STEPS = [(1, 1), (1, 2), (2, 1)]


def _torch():
    import torch
    return torch


def _resolve_use_gpu(use_gpu):
    """Parflex_gpu.py:391-392 / :617 / :691 semantics, minus the CPU fallback."""
    if use_gpu is None:
        use_gpu = os.environ.get("PARFLEX_DISABLE_GPU", "").strip() == ""
    if not use_gpu:
        raise RuntimeError(
            "parflexdtw_amd runs ParFlex on the GPU only; use_gpu=False (or "
            "PARFLEX_DISABLE_GPU) asks for the reference's CPU implementation")
    if not is_gpu_available():
        raise RuntimeError("no ROCm GPU is available (or libparflex_hip.so failed to load)")
    return True


def _sync():
    _torch().cuda.synchronize()


# --------------------------------------------------------------------------
# small helpers of Parflex.py kept for API compatibility
# --------------------------------------------------------------------------

def edge_slot_to_local(edge_type, position, chunk_shape):
    """Parflex.py:139-147."""
    if edge_type == 0:
        return chunk_shape[0] - 1, position
    return position, chunk_shape[1] - 1


def local_to_global_cell(chunk_i, chunk_j, local_row, local_col, chunks_dict):
    """Parflex.py:482-489."""
    s1, _, s2, _ = chunks_dict[(chunk_i, chunk_j)]["bounds"]
    return s1 + local_row, s2 + local_col


def global_cell_to_prev_tile_edge(global_row, global_col, prev_i, prev_j, chunks_dict, L):
    """Parflex.py:492-506."""
    s1, _, s2, _ = chunks_dict[(prev_i, prev_j)]["bounds"]
    lr, lc = global_row - s1, global_col - s2
    shape = chunks_dict[(prev_i, prev_j)]["shape"]
    if lr == shape[0] - 1:
        return 0, lc
    if lc == shape[1] - 1:
        return 1, lr
    raise ValueError(f"({lr},{lc}) not on edge of prev chunk")


def _edge_starts_from_S(S):
    """Parflex.py:347-373: start sets from a tile's last row and column."""
    S = np.asarray(S)
    vals = np.concatenate([S[-1, :], S[:, -1]]).astype(np.int64)
    bot = set(int(v) for v in vals[vals > 0])
    left = set(int(-v) for v in vals[vals <= 0])
    return bot, left


def build_tiled_metadata(chunks_dict, L, n_chunks_1, n_chunks_2, C, stage1_params=None):
    """Parflex.py:1360-1416."""
    if C is None:
        L1 = max(ch["bounds"][1] for ch in chunks_dict.values())
        L2 = max(ch["bounds"][3] for ch in chunks_dict.values())
    else:
        L1, L2 = C.shape
    blocks = []
    for (i, j), ch in chunks_dict.items():
        s1, e1, s2, e2 = ch["bounds"]
        if np.asarray(ch["wp"]).size == 0:
            continue
        blocks.append({"bi": i, "bj": j, "rows": (int(s1), int(e1)), "cols": (int(s2), int(e2)),
                       "Ck_shape": tuple(ch["shape"]), "S_single": ch["S"]})
    if stage1_params is None:
        stage1_params = {"steps": np.array([[1, 1], [1, 2], [2, 1]], dtype=int),
                         "weights": np.array([1.5, 3.0, 3.0], dtype=float), "buffer": 1.0}
    return {"C_shape": (L1, L2), "L_block": L, "hop": L - 1, "n_row": n_chunks_1,
            "n_col": n_chunks_2, "blocks": blocks, "C": C, "stage1_params": stage1_params}


def _write_rows(profile_dir, name, header, rows):
    if profile_dir is None:
        return
    os.makedirs(profile_dir, exist_ok=True)
    with open(os.path.join(profile_dir, name), "w", newline="") as fp:
        w = csv.writer(fp)
        w.writerow(header)
        w.writerows(rows)


# --------------------------------------------------------------------------
# stage 1 with full per-tile outputs (API parity path)
# --------------------------------------------------------------------------

def run_flexdtw_on_tiles(C, L, steps=None, weights=None, buffer=1, profile_dir="Profiling_results",
                         warn_slow_chunks=False, slow_chunk_seconds=1.0, use_gpu=None, C_dev=None):
    """Stage 1 with the reference's per-tile dictionary (Parflex_gpu.py:342-607).

    Each tile's full D, S (=P), B come back to the host, as in the reference;
    parflex() does not use this path (it keeps tile interiors on the device)."""
    steps = _DEFAULT_STEPS if steps is None else steps
    weights = [2, 3, 3] if weights is None else weights
    _resolve_use_gpu(use_gpu)
    if C is None:
        if C_dev is None:
            raise ValueError("run_flexdtw_on_tiles: provide NumPy C and/or C_dev (both None)")
        L1, L2 = int(C_dev.shape[0]), int(C_dev.shape[1])
    else:
        L1, L2 = C.shape
        if C_dev is not None and tuple(C.shape) != tuple(C_dev.shape):
            raise ValueError("C_dev shape must match C")
    dev_buf = C_dev if C_dev is not None else cost_matrix_to_gpu(C)
    hop = L - 1
    n1, n2 = tile_grid(L1, L2, L)
    chunks = {}
    rows = []
    for i in range(n1):
        for j in range(n2):
            t0 = time.perf_counter()
            s1, e1, s2, e2 = tile_bounds(i, j, L1, L2, L)
            bc, wp, D, P, B, debug = flexdtw_chunk_from_global_C(
                dev_buf, int(L2), s1, e1, s2, e2, steps, weights, buffer=buffer)
            bot, left = _edge_starts_from_S(P)
            chunks[(i, j)] = {
                "C": None if C is None else C[s1:e1, s2:e2], "D": D, "S": P, "B": B,
                "debug": debug, "best_cost": bc, "wp": wp, "bounds": (s1, e1, s2, e2),
                "hop": (hop if e1 < L1 else L1 - s1, hop if e2 < L2 else L2 - s2),
                "shape": (e1 - s1, e2 - s2), "starts_bot_edge": bot, "starts_left_edge": left,
            }
            el = time.perf_counter() - t0
            rows.append([i, j, t0, t0 + el, el])
            if warn_slow_chunks and el > slow_chunk_seconds:
                print(f"  Warning, chunk ({i},{j}) took {el} seconds")
    _write_rows(profile_dir, "chunk_flexdtw.csv",
                ["chunk_i", "chunk_j", "start_time", "end_time", "elapsed_seconds"], rows)
    return chunks, L, n1, n2


# --------------------------------------------------------------------------
# device pipeline
# --------------------------------------------------------------------------

class _DeviceTiles:
    """Stage-1 state kept on the GPU between tiled_stage1_from_features and
    run_stage2_from_tiled: the plan, workspace, step set and C.  L_api is the
    caller's tile length; plan.L is the one the device runs with (see
    _run_L)."""

    def __init__(self, C_dev, plan, sw, ws, L_api=None):
        self.C_dev, self.plan, self.sw, self.ws = C_dev, plan, sw, ws
        self.L_api = plan.L if L_api is None else int(L_api)


def _run_L(L, L1, L2):
    """Tile length the device runs with.  A tile length of at least
    max(L1, L2) gives one tile covering the whole matrix (Parflex.py:388-391);
    every tile of length max(L1, L2) gives the same tile, segment numbering
    (g // L is 0 either way) and result, with stage-2 arrays of that length
    instead of L (padded back with +inf, their value in the reference)."""
    return min(int(L), max(int(L1), int(L2)))


def _pad_L(a, L_api):
    """Stage-2 arrays (n1, n2, 2, L) padded with +inf to the caller's L."""
    if a.shape[-1] == L_api:
        return a
    out = np.full(a.shape[:-1] + (L_api,), np.inf)
    out[..., :a.shape[-1]] = a
    return out


def _stage1_device(C_dev, L, steps, weights, beta):
    L1, L2 = int(C_dev.shape[0]), int(C_dev.shape[1])
    plan = BatchPlan([(L1, L2)], _run_L(L, L1, L2), beta=float(beta))
    sw = StepWeights(steps, weights)
    ws = plan.allocate(C_dev.device)
    engine.stage1(C_dev, plan, sw, ws)
    return _DeviceTiles(C_dev, plan, sw, ws, L_api=L)


def _set_beta(dt, beta):
    L1, L2 = int(dt.plan.pairs[0, 0]), int(dt.plan.pairs[0, 1])
    buf = int(buffer_global(L1, L2, beta))
    if dt.plan.pairs[0, 7] != buf:
        dt.plan.pairs[0, 7] = buf
        dt.ws["pairs"].copy_(_torch().from_numpy(dt.plan.pairs))


def parflex(C, steps, weights, beta, L=None, profile_dir="Profiling_results",
            backtrace_segments=False, warn_slow_chunks=False, slow_chunk_seconds=1.0,
            use_gpu=None):
    """End-to-end ParFlex on a cost matrix (Parflex_gpu.py:656-854).

    C: host array (L1, L2) or a ROCm tensor.  A device tensor of dtype
    bfloat16 or float32 selects the precision form (FP32 stage-1 DP, FP64
    stage 2); anything else runs in FP64 and matches the reference exactly.
    Returns (best_cost, wp) with wp int64 of shape (2, n); raises ValueError
    when no endpoint exists."""
    _resolve_use_gpu(use_gpu)
    L = DEFAULT_CHUNK_LENGTH if L is None else int(L)
    t_all = time.perf_counter()
    t0 = time.perf_counter()
    torch = _torch()
    if (isinstance(C, torch.Tensor) and C.is_cuda and C.dim() == 2
            and C.dtype in (torch.bfloat16, torch.float32)):
        C_dev = C.contiguous()
    else:
        C_dev = cost_matrix_to_gpu(C)
    if profile_dir is not None:
        _sync()
    t_h2d = time.perf_counter() - t0
    t0 = time.perf_counter()
    dt = _stage1_device(C_dev, L, steps, weights, beta)
    if profile_dir is not None:
        _sync()
    t_s1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    engine.stage2(C_dev, dt.plan, dt.sw, dt.ws)
    res = engine.collect(dt.plan, dt.ws)
    t_s2 = time.perf_counter() - t0
    if profile_dir is not None:
        total = time.perf_counter() - t_all
        _write_rows(profile_dir, "parflex_end_to_end_phases.csv",
                    ["phase", "elapsed_seconds", "had_device_C"],
                    [["gpu_host_to_device_C", t_h2d, 1],
                     ["stage1_chunk_flexdtw_all_tiles", t_s1, 1],
                     ["stage2_propagate_sync_scan_stitch", t_s2, 1],
                     ["parflex_total_runtime", total, 1]])
    if not res.found[0]:
        raise ValueError(NO_ENDPOINT_MSG)
    wp = res.paths[0]
    if wp.size == 0:
        wp = np.array([[], []], dtype=np.int64)
    return float(res.best_cost[0]), wp


def tiled_stage1_from_features(F1, F2, steps=None, weights=None, beta=0.1, L=None, use_gpu=None,
                               cost_dtype=None):
    """Features -> cost matrix on the GPU -> stage 1 (Parflex_gpu.py:610-653).

    Returns (None, tiled_result); tiled_result['C_dev'] holds the device cost
    matrix and tiled_result['chunks_dict'] materialises a tile's full D/S/B on
    first access (they are not kept on the device by stage 1).  cost_dtype
    torch.bfloat16 or torch.float32 selects the precision form (MFMA cost
    build, FP32 DP); the default FP64 matches the reference exactly."""
    _resolve_use_gpu(use_gpu)
    L = DEFAULT_CHUNK_LENGTH if L is None else int(L)
    steps = np.array([[1, 1], [1, 2], [2, 1]]) if steps is None else steps
    weights = np.array([1.25, 3.0, 3.0]) if weights is None else weights
    C_dev = cost_matrix_from_features_gpu(F1, F2, dtype=cost_dtype)
    dt = _stage1_device(C_dev, L, steps, weights, beta)
    n1, n2 = dt.plan.grids[0]
    L1, L2 = int(C_dev.shape[0]), int(C_dev.shape[1])
    stage1_params = {"steps": np.array(steps).reshape((-1, 2)), "weights": np.array(weights),
                     "buffer": 1.0}
    chunks = _LazyChunks(C_dev, L1, L2, L, n1, n2, steps, weights)
    chunks.device = dt
    tiled = {"C_shape": (L1, L2), "L_block": L, "hop": L - 1, "n_row": n1, "n_col": n2,
             "blocks": [], "C": None, "stage1_params": stage1_params, "chunks_dict": chunks,
             "C_dev": C_dev, "_device": dt}
    return None, tiled


class _LazyChunks(dict):
    """chunks_dict whose per-tile entries are computed on first access with
    the full-output tile kernel (the reference keeps them all on the host)."""

    def __init__(self, C_dev, L1, L2, L, n1, n2, steps, weights):
        super().__init__()
        self.device = None  # the stage-1 state on the GPU (_DeviceTiles)
        self._args = (C_dev, L1, L2, L, steps, weights)
        self._keys = [(i, j) for i in range(n1) for j in range(n2)]

    def _fill(self, key):
        C_dev, L1, L2, L, steps, weights = self._args
        i, j = key
        s1, e1, s2, e2 = tile_bounds(i, j, L1, L2, L)
        bc, wp, D, P, B, debug = flexdtw_chunk_from_global_C(C_dev, L2, s1, e1, s2, e2, steps,
                                                             weights, buffer=1)
        bot, left = _edge_starts_from_S(P)
        hop = L - 1
        dict.__setitem__(self, key, {
            "C": None, "D": D, "S": P, "B": B, "debug": debug, "best_cost": bc, "wp": wp,
            "bounds": (s1, e1, s2, e2),
            "hop": (hop if e1 < L1 else L1 - s1, hop if e2 < L2 else L2 - s2),
            "shape": (e1 - s1, e2 - s2), "starts_bot_edge": bot, "starts_left_edge": left})

    def __getitem__(self, key):
        if not dict.__contains__(self, key):
            if key not in self._keys:
                raise KeyError(key)
            self._fill(key)
        return dict.__getitem__(self, key)

    def __contains__(self, key):
        return key in self._keys

    def __iter__(self):
        return iter(self._keys)

    def __len__(self):
        return len(self._keys)

    def keys(self):
        return list(self._keys)

    def items(self):
        return [(k, self[k]) for k in self._keys]

    def values(self):
        return [self[k] for k in self._keys]


def _stage2_dict(dt, backtrace_segments, arrays=None):
    """Assemble the stage-2 result dict of Parflex.py:1342-1356 from device
    buffers; arrays = (D_arr, L_arr, edge_lens) of the caller are returned as
    they are (stage2_scan_and_stitch returns its inputs)."""
    plan, ws, sw = dt.plan, dt.ws, dt.sw
    L = plan.L
    n1, n2 = plan.grids[0]
    L1, L2 = int(plan.pairs[0, 0]), int(plan.pairs[0, 1])
    res = engine.collect(plan, ws)
    if not res.found[0]:
        raise ValueError(NO_ENDPOINT_MSG)
    if arrays is None:
        D_arr = _pad_L(ws["D_arr"].cpu().numpy().reshape(n1, n2, 2, L), dt.L_api)
        L_arr = _pad_L(ws["L_arr"].cpu().numpy().reshape(n1, n2, 2, L), dt.L_api)
        edge_lens = _edge_lens(plan)
    else:
        D_arr, L_arr, edge_lens = arrays
    candD = ws["cand_D"].cpu().numpy()
    candL = ws["cand_L"].cpu().numpy()
    top_D, top_L = candD[:L2].copy(), candL[:L2].copy()
    right_D, right_L = candD[L2:L2 + L1].copy(), candL[L2:L2 + L1].copy()
    top_norm = np.full(L2, np.nan)
    right_norm = np.full(L1, np.nan)
    m = (top_L > 0) & np.isfinite(top_D)
    top_norm[m] = top_D[m] / top_L[m]
    m = (right_L > 0) & np.isfinite(right_D)
    right_norm[m] = right_D[m] / right_L[m]
    segf = ws["seg_f"].cpu().numpy().reshape(-1, 4)
    segi = ws["seg_i"].cpu().numpy().reshape(-1, 8)
    max_segs = int(plan.pairs[0, 10])
    n_segs = (-(-L2 // L), -(-L1 // L))
    names = ("top", "right")
    best_per_segment = {}
    for side in range(2):
        for s in range(n_segs[side]):
            k = side * max_segs + s
            if segi[k, 6] != 1:
                continue
            key = (names[side], s)
            best_per_segment[key] = {
                "chunk_i": int(segi[k, 0]), "chunk_j": int(segi[k, 1]), "edge": int(segi[k, 2]),
                "idx": int(segi[k, 3]), "norm_cost": float(segf[k, 0]),
                "global_coord": (int(segi[k, 4]), int(segi[k, 5])), "segment": key}
    paths_per_segment = {}
    if backtrace_segments and best_per_segment:
        keys = list(best_per_segment)
        req = [[0, best_per_segment[k]["chunk_i"], best_per_segment[k]["chunk_j"],
                best_per_segment[k]["edge"], best_per_segment[k]["idx"]] for k in keys]
        paths = engine.stitch_requests(plan, sw, ws, req)
        for k, p in zip(keys, paths):
            paths_per_segment[k] = {"endpoint": best_per_segment[k],
                                    "path": [(int(a), int(b)) for a, b in p]}
    stitched = res.paths[0].T.copy() if res.paths[0].size else np.zeros((0, 2), dtype=int)
    be = res.best_end[0]
    return {
        "D_arr": D_arr, "L_arr": L_arr, "edge_lens": edge_lens,
        "best_cost": float(res.best_cost[0]), "best_end": tuple(int(v) for v in be),
        "stitched_wp": stitched, "n_row": n1, "n_col": n2,
        "edge_summary": {"top": {"D": top_D, "L": top_L, "norm": top_norm},
                         "right": {"D": right_D, "L": right_L, "norm": right_norm}},
        "best_per_segment": best_per_segment, "paths_per_segment": paths_per_segment,
    }


def _edge_lens(plan):
    """edge_lens (n1, n2, 2) of Parflex.py:509-564: columns, then rows."""
    n1, n2 = plan.grids[0]
    return plan.tiles[:, [3, 2]].reshape(n1, n2, 2).copy()


def _pack_B(B):
    """Full int8 back-pointers (R, W) -> 2-bit words, 16 rows per uint32 per
    column (the stage-1 layout of include/parflex_hip.h)."""
    B = np.asarray(B)
    R, W = B.shape
    nb = -(-R // 16)
    full = np.zeros((nb * 16, W), np.uint32)
    full[:R] = B.astype(np.uint32) & 3
    full = full.reshape(nb, 16, W)
    words = np.zeros((nb, W), np.uint32)
    for k in range(16):
        words |= full[:, k, :] << np.uint32(2 * k)
    return words.reshape(-1)


def _device_from_host_chunks(chunks_dict, L, n1, n2, steps, C=None, C_dev=None):
    """Stage-2 inputs on the device from a host chunks_dict (per-tile D, S, B,
    C and bounds, the dict of run_flexdtw_on_tiles, Parflex.py:446-454): edge
    D and P of the last row and column, C of the first row and column, and
    the packed back-pointers."""
    torch = _torch()
    L = int(L)
    keys = [(i, j) for i in range(n1) for j in range(n2)]
    for k in keys:
        if k not in chunks_dict:
            raise ValueError(f"chunks_dict has no tile {k}")
    L1 = max(int(chunks_dict[k]["bounds"][1]) for k in keys)
    L2 = max(int(chunks_dict[k]["bounds"][3]) for k in keys)
    L_api, L = L, _run_L(L, L1, L2)
    plan = BatchPlan([(L1, L2)], L)
    if tuple(plan.grids[0]) != (int(n1), int(n2)):
        raise ValueError(f"tile grid {plan.grids[0]} of the chunk bounds differs from "
                         f"({n1}, {n2})")
    T = plan.ntiles
    eD = np.zeros(T * 2 * L)
    eP = np.zeros(T * 2 * L, np.int32)
    eC = np.zeros(T * 2 * L)
    Bp = np.zeros(max(plan.b_words, 1), np.uint32)
    C_host = None if C is None else np.asarray(C)
    for t, (i, j) in enumerate(keys):
        ch = chunks_dict[(i, j)]
        s1, e1, s2, e2 = (int(v) for v in ch["bounds"])
        if (s1, e1, s2, e2) != tile_bounds(i, j, L1, L2, L):
            raise ValueError(f"tile {(i, j)}: bounds {ch['bounds']} are not the tiling of L={L}")
        R, W = e1 - s1, e2 - s2
        D = np.asarray(ch["D"], np.float64)
        S = np.asarray(ch["S"])
        o = t * 2 * L
        eD[o:o + W] = D[R - 1, :]
        eD[o + L:o + L + R] = D[:, W - 1]
        eP[o:o + W] = S[R - 1, :]
        eP[o + L:o + L + R] = S[:, W - 1]
        Ct = ch.get("C")
        if Ct is not None:
            Ct = np.asarray(Ct)
            eC[o:o + W] = Ct[0, :]
            eC[o + L:o + L + R] = Ct[:, 0]
        elif C_host is not None:
            eC[o:o + W] = C_host[s1, s2:e2]
            eC[o + L:o + L + R] = C_host[s1:e1, s2]
        elif C_dev is not None:
            eC[o:o + W] = C_dev[s1, s2:e2].double().cpu().numpy()
            eC[o + L:o + L + R] = C_dev[s1:e1, s2].double().cpu().numpy()
        else:
            raise ValueError("stage 2 needs the cost of each tile's first row and column: "
                             "the chunks hold no C and no C or C_dev was given")
        b0 = int(plan.tiles[t, 4])
        Bp[b0:b0 + (-(-R // 16)) * W] = _pack_B(ch["B"])
    dev = torch.device("cuda", torch.cuda.current_device())
    ws = plan.allocate(dev)
    ws["edge_D"].copy_(torch.from_numpy(eD))
    ws["edge_P"].copy_(torch.from_numpy(eP))
    ws["edge_C"].copy_(torch.from_numpy(eC))
    ws["Bp"].copy_(torch.from_numpy(Bp.view(np.int32)))
    weights = np.ones(len(np.asarray(steps).reshape(-1, 2)))
    return _DeviceTiles(None, plan, StepWeights(steps, weights), ws, L_api=L_api)


def _device_for(chunks_dict, L, n1, n2, tiled_result=None, C_dev=None):
    """The device stage-1 state behind chunks_dict: the one kept by
    tiled_stage1_from_features, the one built earlier for the same dict (kept
    in tiled_result), or a new one built from the host arrays."""
    dt = getattr(chunks_dict, "device", None)
    if dt is not None:
        if dt.L_api != int(L) or tuple(dt.plan.grids[0]) != (int(n1), int(n2)):
            raise ValueError("L or the tile grid differ from the stage-1 run of chunks_dict")
        return dt
    cached = None if tiled_result is None else tiled_result.get("_device_chunks")
    if cached is not None and cached[0] is chunks_dict and cached[1].L_api == int(L):
        return cached[1]
    steps = _DEFAULT_STEPS
    C = None
    if tiled_result is not None:
        steps = tiled_result.get("stage1_params", {}).get("steps", steps)
        C = tiled_result.get("C")
        if C_dev is None:
            C_dev = tiled_result.get("C_dev")
    dt = _device_from_host_chunks(chunks_dict, L, int(n1), int(n2), steps, C=C, C_dev=C_dev)
    if tiled_result is not None:
        tiled_result["_device_chunks"] = (chunks_dict, dt)
    return dt


def propagate_tile_edge_costs(chunks_dict, L, num_chunks_1, num_chunks_2, buffer_param=0.1,
                              profile_dir="Profiling_results", C_dev=None):
    """Stage-2 propagation on the device (Parflex_gpu.py:247-339,
    Parflex.py:862-912): edge data, valid mask, init and fill.  Returns host
    (D_arr, L_arr, edge_lens) before the overlap sync.  buffer_param is
    unused, as in the reference.  The device runs init and fill as one sweep
    over the tile anti-diagonals, so initialize_chunks.csv and
    dp_fill_chunks.csv both record that sweep."""
    _resolve_use_gpu(None)
    dt = _device_for(chunks_dict, L, num_chunks_1, num_chunks_2, C_dev=C_dev)
    plan, ws = dt.plan, dt.ws
    n1, n2 = plan.grids[0]
    t0 = time.perf_counter()
    engine.stage2_phases(plan, dt.sw, ws, PF_S2_PROPAGATE)
    D_arr = _pad_L(ws["D_arr"].cpu().numpy().reshape(n1, n2, 2, plan.L).copy(), dt.L_api)
    L_arr = _pad_L(ws["L_arr"].cpu().numpy().reshape(n1, n2, 2, plan.L).copy(), dt.L_api)
    t1 = time.perf_counter()
    row = [["all", "all", t0, t1, t1 - t0]]
    hdr = ["chunk_i", "chunk_j", "start_time", "end_time", "elapsed_seconds"]
    _write_rows(profile_dir, "initialize_chunks.csv", hdr, row)
    _write_rows(profile_dir, "dp_fill_chunks.csv", hdr, row)
    return D_arr, L_arr, _edge_lens(plan)


def sync_tile_overlap_edges(D_arr, L_arr, edge_lens, n_chunks_1, n_chunks_2):
    """The three overlap passes of Parflex.py:919-969 on the device, written
    back into D_arr and L_arr in place; returns (D_arr, L_arr)."""
    _resolve_use_gpu(None)
    torch = _torch()
    n1, n2 = int(n_chunks_1), int(n_chunks_2)
    D = np.asarray(D_arr)
    La = np.asarray(L_arr)
    el = np.asarray(edge_lens).astype(np.int64)
    if D.ndim != 4 or D.shape[:3] != (n1, n2, 2) or La.shape != D.shape:
        raise ValueError(f"D_arr and L_arr must have shape ({n1}, {n2}, 2, L)")
    if el.shape != (n1, n2, 2):
        raise ValueError(f"edge_lens must have shape ({n1}, {n2}, 2)")
    L = int(D.shape[3])
    if np.any(el < 0) or np.any(el > L) or np.any(el == 1):
        raise ValueError("edge lengths must be 0 or in [2, L]")
    tiles = np.zeros((n1 * n2, 8), np.int64)
    tiles[:, 2] = el[..., 1].reshape(-1)
    tiles[:, 3] = el[..., 0].reshape(-1)
    pairs = np.zeros((1, 16), np.int64)
    pairs[0, 2], pairs[0, 3] = n1, n2
    dev = torch.device("cuda", torch.cuda.current_device())
    ws = {"tiles": torch.from_numpy(tiles).to(dev), "pairs": torch.from_numpy(pairs).to(dev),
          "D_arr": torch.from_numpy(np.ascontiguousarray(D, np.float64).reshape(-1)).to(dev),
          "L_arr": torch.from_numpy(np.ascontiguousarray(La, np.float64).reshape(-1)).to(dev)}
    grid = types.SimpleNamespace(npairs=1, ntiles=n1 * n2, L=L, max_diag=1, max_tile_diag=1,
                                 max_tiles_pair=n1 * n2, max_segs=1)
    engine.stage2_phases(grid, StepWeights(_DEFAULT_STEPS, [1, 1, 1]), ws, PF_S2_SYNC)
    D_out = ws["D_arr"].cpu().numpy().reshape(D.shape)
    L_out = ws["L_arr"].cpu().numpy().reshape(D.shape)
    if isinstance(D_arr, np.ndarray) and isinstance(L_arr, np.ndarray):
        D_arr[...] = D_out
        L_arr[...] = L_out
        return D_arr, L_arr
    return D_out, L_out


def stage2_scan_and_stitch(tiled_result, chunks_dict, D_arr, L_arr, edge_lens, L1, L2, L_block,
                           buffer_stage2=200, top_k=1, profile_dir="Profiling_results",
                           backtrace_segments=False):
    """Boundary scan, segment winners and stitched backtrace on the device
    (Parflex.py:1229-1356) from the caller's (synchronised) D_arr and L_arr.
    Returns the reference's dict; raises ValueError when no endpoint exists."""
    _resolve_use_gpu(None)
    torch = _torch()
    n1, n2 = int(tiled_result["n_row"]), int(tiled_result["n_col"])
    dt = _device_for(chunks_dict, L_block, n1, n2, tiled_result=tiled_result)
    plan, ws = dt.plan, dt.ws
    if (int(plan.pairs[0, 0]), int(plan.pairs[0, 1])) != (int(L1), int(L2)):
        raise ValueError(f"(L1, L2) = ({L1}, {L2}) differs from the tiles' extent "
                         f"({plan.pairs[0, 0]}, {plan.pairs[0, 1]})")
    D = np.asarray(D_arr, np.float64)
    La = np.asarray(L_arr, np.float64)
    if D.shape != (n1, n2, 2, dt.L_api) or La.shape != D.shape:
        raise ValueError(f"D_arr and L_arr must have shape ({n1}, {n2}, 2, {dt.L_api})")
    D = np.ascontiguousarray(D[..., :plan.L])
    La = np.ascontiguousarray(La[..., :plan.L])
    t0 = time.perf_counter()
    buf = int(buffer_stage2)
    if plan.pairs[0, 7] != buf:
        plan.pairs[0, 7] = buf
        ws["pairs"].copy_(torch.from_numpy(plan.pairs))
    ws["D_arr"].copy_(torch.from_numpy(D.reshape(-1)))
    ws["L_arr"].copy_(torch.from_numpy(La.reshape(-1)))
    engine.stage2_phases(plan, dt.sw, ws, PF_S2_SCAN | PF_S2_STITCH)
    out = _stage2_dict(dt, backtrace_segments, arrays=(D_arr, L_arr, edge_lens))
    t1 = time.perf_counter()
    _write_rows(profile_dir, "stage_2_backtrace_compatible.csv",
                ["phase", "start_time", "end_time", "elapsed_seconds"],
                [["edge_scan_and_backtrace_stitch", t0, t1, t1 - t0]])
    return out


def backtrace_and_stitch(tiled_result, chunks_dict, start_i, start_j, start_edge, start_idx):
    """Stitched path from one edge slot across at most 101 tiles
    (Parflex.py:1032-1103), walked on the device.  Returns the list of
    global (row, col) points, first point first."""
    _resolve_use_gpu(None)
    n1, n2 = int(tiled_result["n_row"]), int(tiled_result["n_col"])
    i, j = int(start_i), int(start_j)
    if (i, j) not in chunks_dict or not (0 <= i < n1 and 0 <= j < n2):
        return []
    dt = _device_for(chunks_dict, tiled_result["L_block"], n1, n2, tiled_result=tiled_result)
    e = 0 if int(start_edge) == 0 else 1
    t = dt.plan.tiles[dt.plan.tile_index(0, i, j)]
    n = int(t[3]) if e == 0 else int(t[2])
    idx = int(start_idx)
    if not 0 <= idx < n:
        raise IndexError(f"edge index {idx} outside edge {e} of tile {(i, j)} (length {n})")
    steps = tiled_result.get("stage1_params", {}).get("steps")
    sw = dt.sw
    if steps is not None and not np.array_equal(np.asarray(steps).reshape(-1, 2),
                                                sw.steps_i64):
        sw = StepWeights(steps, np.ones(len(np.asarray(steps).reshape(-1, 2))))
    path = engine.stitch_requests(dt.plan, sw, dt.ws, [[0, i, j, e, idx]])[0]
    return [(int(a), int(b)) for a, b in path]


def run_stage2_from_tiled(tiled_result, C=None, beta=0.1, show_fig=False, top_k=1,
                          profile_dir=None, backtrace_segments=False):
    """Stage 2 on the device (Parflex_gpu.py:857-910); returns the dict of
    Parflex.stage2_scan_and_stitch (Parflex.py:1342-1356)."""
    dt = tiled_result.get("_device")
    if dt is None:
        C_dev = tiled_result.get("C_dev")
        if C_dev is None:
            if C is None:
                C = tiled_result.get("C")
            if C is None:
                raise ValueError("run_stage2_from_tiled needs C, C_dev or a device stage 1")
            C_dev = cost_matrix_to_gpu(C)
        p = tiled_result["stage1_params"]
        dt = _stage1_device(C_dev, tiled_result["L_block"], p["steps"], p["weights"], beta)
        tiled_result["_device"] = dt
    _set_beta(dt, beta)
    t0 = time.perf_counter()
    engine.stage2(dt.C_dev, dt.plan, dt.sw, dt.ws)
    out = _stage2_dict(dt, backtrace_segments)
    if profile_dir is not None:
        _write_rows(profile_dir, "stage_2_backtrace_compatible.csv",
                    ["phase", "elapsed_seconds"], [["stage2_total", time.perf_counter() - t0]])
    return out
