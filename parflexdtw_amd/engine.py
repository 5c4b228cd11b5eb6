"""Device pipeline: stage 1 (all tiles, one launch) and stage 2 (on device).

This is the layer the drop-in modules (gpu_flexdtw, Parflex_gpu) and the
benchmark call.  Inputs are torch tensors resident on the GPU; only the
stitched paths and best costs are copied back.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import PF_BF16, PF_F32, PF_F64, check, lib
from ._plan import MAX_STITCH_TILES, BatchPlan

__all__ = ["StepWeights", "stage1", "stage2", "stage2_phases", "run_batch", "BatchResult",
           "dtype_code", "NO_ENDPOINT_MSG", "BatchRunner", "autotune_stage1", "stage1_config",
           "tune_batch"]

NO_ENDPOINT_MSG = "Stage 2: No valid endpoint found on global top/right edges."


def _torch():
    import torch
    return torch


def dtype_code(t):
    torch = _torch()
    if t.dtype == torch.float64:
        return PF_F64
    if t.dtype == torch.float32:
        return PF_F32
    if t.dtype == torch.bfloat16:
        return PF_BF16
    raise ValueError(f"unsupported cost-matrix dtype {t.dtype}")


class StepWeights:
    """Host copies of the step set in the C-ABI layout (int32 pairs, f64)."""

    def __init__(self, steps, weights):
        s = np.asarray(steps, dtype=np.int64).reshape(-1, 2)
        w = np.asarray(weights, dtype=np.float64).reshape(-1)
        if s.shape[0] != w.shape[0]:
            raise ValueError("steps and weights must have the same length")
        self.steps = np.ascontiguousarray(s.astype(np.int32))
        self.weights = np.ascontiguousarray(w)
        self.K = int(s.shape[0])
        check(lib().pf_check_steps(self.steps.ctypes.data, self.weights.ctypes.data, self.K),
              "step set")
        self.steps_i64 = s

    @property
    def sp(self):
        return self.steps.ctypes.data

    @property
    def wp(self):
        return self.weights.ctypes.data


def _stream(stream):
    torch = _torch()
    if stream is None:
        stream = torch.cuda.current_stream()
    return stream.cuda_stream


def _cpl_for(width, strips, waves):
    for cpl in (2, 4, 8):
        if strips * waves * 64 * cpl >= width:
            return cpl
    return 0


def stage1_config(plan, n_cu=256):
    """Launch shape of the stage-1 DP: (strips per tile, waves per workgroup,
    rows per exchange block).  Few tiles are cut into column strips so that
    the grid covers the machine; many tiles run one workgroup each.
    PARFLEX_S1_STRIPS / PARFLEX_S1_WAVES / PARFLEX_S1_G override the choice."""
    L = plan.L
    env = os.environ
    if "PARFLEX_S1_STRIPS" in env or "PARFLEX_S1_WAVES" in env:
        S = int(env.get("PARFLEX_S1_STRIPS", "1"))
        nw = int(env.get("PARFLEX_S1_WAVES", "0"))
        if nw <= 0:
            nw = max(1, -(-L // (S * 128)))
        G = int(env.get("PARFLEX_S1_G", "16"))
        return S, nw, G
    # Workgroups of up to 4 waves, 2 columns per lane (512-column strips):
    # the best measured shape for both one 16000 x 18209 pair at L = 4000
    # (8 strips per tile, exchange blocks of 16 rows) and 1195 tiles at
    # L = 500 (one workgroup per tile).
    nw = max(1, min(4, -(-L // 128)))
    S = -(-L // (nw * 128))
    return S, nw, int(env.get("PARFLEX_S1_G", "16"))


def stage1(C_dev, plan, sw, ws, stream=None, config=None, feat=None):
    """Per-tile DP for every tile of the plan: packed back-pointers plus edge
    D/P (replaces the tile loop at Parflex_gpu.py:461-563).  With feat =
    (F1b, F2b, Kp) the cost is computed inside the sweep from BF16 features
    (pf_stage1_features; ws["tiles"] columns 0/1 then hold frame indices)
    and C_dev is not used."""
    S, nw, G = config or ws.get("s1_config") or stage1_config(plan)
    torch = _torch()
    dev = ws["tiles"].device
    code = PF_BF16 if feat is not None else dtype_code(C_dev)
    if S > 1:
        nf, nx = ctypes.c_int64(), ctypes.c_int64()
        check(lib().pf_stage1_workspace(code, plan.ntiles, plan.L, S,
                                        ctypes.byref(nf), ctypes.byref(nx)), "pf_stage1_workspace")
        need_f = nf.value
        need_x = -(-nx.value // 8)
        if ws.get("s1_flags") is None or ws["s1_flags"].numel() < need_f:
            ws["s1_flags"] = torch.empty(need_f, dtype=torch.int32, device=dev)
        if ws.get("s1_xch") is None or ws["s1_xch"].numel() < need_x:
            ws["s1_xch"] = torch.empty(need_x, dtype=torch.float64, device=dev)
        fl, xc = ws["s1_flags"].data_ptr(), ws["s1_xch"].data_ptr()
    else:
        fl = xc = None
    if feat is not None:
        F1b, F2b, Kp = feat
        rc = lib().pf_stage1_features(F1b.data_ptr(), F2b.data_ptr(), int(Kp),
                                      ws["tiles"].data_ptr(), plan.ntiles, plan.L, sw.sp, sw.wp,
                                      sw.K, S, nw, G, fl, xc, ws["Bp"].data_ptr(),
                                      ws["edge_D"].data_ptr(), ws["edge_P"].data_ptr(),
                                      ws["edge_C"].data_ptr(), _stream(stream))
        check(rc, "pf_stage1_features")
        return
    rc = lib().pf_stage1(dtype_code(C_dev), C_dev.data_ptr(), ws["tiles"].data_ptr(),
                         plan.ntiles, plan.L, sw.sp, sw.wp, sw.K, S, nw, G, fl, xc,
                         ws["Bp"].data_ptr(), ws["edge_D"].data_ptr(), ws["edge_P"].data_ptr(),
                         ws["edge_C"].data_ptr(), _stream(stream))
    check(rc, "pf_stage1")


def device_status(stream=None):
    """Synchronise the stream and raise ParflexError if a stage-1 wait gave
    up since the last check (pf_device_status; the results are then not
    valid)."""
    check(lib().pf_device_status(_stream(stream)), "stage 1")


def _dag_tables(plan, ws):
    """Device tables of the stage-2 tile DAG over every tile of the plan
    (all with stage 1): the diagonal-ordered tile list, its per-diagonal
    table and the flags / ticket workspace, made once per workspace."""
    t = ws.get("s2_dag")
    if t is None:
        torch = _torch()
        dev = ws["tiles"].device
        lst, h_diag = diagonal_lists(plan, np.ones(plan.ntiles, bool))
        t = {"list": torch.from_numpy(lst).to(dev), "diag": torch.from_numpy(h_diag).to(dev),
             "work": torch.empty(plan.ntiles + 1, dtype=torch.int32, device=dev)}
        ws["s2_dag"] = t
    return t


def stage2(C_dev, plan, sw, ws, stream=None, dag=False):
    """Propagation, sync, scan and stitch on the device (Parflex.py:862-1356).
    With dag=True (every tile of the plan with stage 1) the propagation runs
    as one tile-DAG launch (pf_stage2_propagate_dag) instead of a launch per
    tile diagonal; the results are the same."""
    if dag:
        t = _dag_tables(plan, ws)
        _propagate_dag(plan, ws, t["list"], t["diag"], t["work"], stream)
        stage2_phases(plan, sw, ws, _lib.PF_S2_SYNC | _lib.PF_S2_SCAN | _lib.PF_S2_STITCH, stream)
        return
    rc = lib().pf_stage2(
        ws["edge_C"].data_ptr(), ws["pairs"].data_ptr(), plan.npairs,
        ws["tiles"].data_ptr(), plan.ntiles, plan.L, plan.max_diag, plan.max_tile_diag,
        plan.max_tiles_pair, plan.max_segs, sw.sp, sw.K, ws["Bp"].data_ptr(),
        ws["edge_D"].data_ptr(), ws["edge_P"].data_ptr(), ws["D_arr"].data_ptr(),
        ws["L_arr"].data_ptr(), ws["olap"].data_ptr(), ws["bits"].data_ptr(),
        ws["cand_D"].data_ptr(), ws["cand_L"].data_ptr(), ws["seg_f"].data_ptr(),
        ws["seg_i"].data_ptr(), ws["best_f"].data_ptr(), ws["best_i"].data_ptr(),
        ws["chain"].data_ptr(), ws["pts"].data_ptr(), ws["path"].data_ptr(),
        ws["path_len"].data_ptr(), _stream(stream))
    check(rc, "pf_stage2")


def stage2_phases(plan, sw, ws, phases, stream=None):
    """Some phases of stage 2 (pf_stage2_phases): _lib.PF_S2_PROPAGATE, _SYNC,
    _SCAN, _STITCH, run in that order; workspace entries that no requested
    phase uses may be missing from ws."""
    def ptr(k):
        t = ws.get(k)
        return None if t is None else t.data_ptr()
    rc = lib().pf_stage2_phases(
        int(phases), ptr("edge_C"), ws["pairs"].data_ptr(), plan.npairs, ws["tiles"].data_ptr(),
        plan.ntiles, plan.L, plan.max_diag, plan.max_tile_diag, plan.max_tiles_pair,
        plan.max_segs, sw.sp, sw.K, ptr("Bp"), ptr("edge_D"), ptr("edge_P"),
        ws["D_arr"].data_ptr(), ws["L_arr"].data_ptr(), ptr("olap"), ptr("bits"), ptr("cand_D"),
        ptr("cand_L"), ptr("seg_f"), ptr("seg_i"), ptr("best_f"), ptr("best_i"), ptr("chain"),
        ptr("pts"), ptr("path"), ptr("path_len"), _stream(stream))
    check(rc, "pf_stage2_phases")


def diagonal_lists(plan, live, cache=None):
    """Tile lists of the tile anti-diagonals (diagonal of tile (i, j) = i + j)
    for pf_stage2_propagate_listed: every tile of the plan once, diagonal by
    diagonal, the tiles with live[t] first.  Returns (list int32, h_diag
    int64 (max_diag, 3): offset, live count, other count).  `cache` (a dict)
    keeps the per-plan diagonal order between calls."""
    if cache is None or "order" not in cache:
        ij = plan.tiles[:, 7]
        d = ((ij >> 32) + (ij & 0xFFFFFFFF)).astype(np.int64)
        c = {"d": d, "order": np.argsort(d, kind="stable"),
             "count": np.bincount(d, minlength=plan.max_diag)}
        if cache is not None:
            cache.update(c)
        cache = c
    d, order, nt = cache["d"], cache["order"], cache["count"]
    live = np.asarray(live, bool)
    nd = plan.max_diag
    nl = np.bincount(d[live], minlength=nd)
    npr = nt - nl
    off = np.concatenate([[0], np.cumsum(nt)[:-1]])
    s_ord = live[order]
    dd = d[order]
    live_start = np.concatenate([[0], np.cumsum(nl)[:-1]])
    prun_start = np.concatenate([[0], np.cumsum(npr)[:-1]])
    live_rank = np.cumsum(s_ord) - 1
    prun_rank = np.cumsum(~s_ord) - 1
    pos = np.where(s_ord, off[dd] + live_rank - live_start[dd],
                   off[dd] + nl[dd] + prun_rank - prun_start[dd])
    lst = np.empty(len(order), np.int32)
    lst[pos] = order
    return lst, np.ascontiguousarray(np.stack([off, nl, npr], axis=1).astype(np.int64))


def stage2_listed(plan, sw, ws, live, stream=None, cache=None, dag=False):
    """stage2 for a batch whose tiles with live[t] False had no stage 1
    (tile table column 0 < 0): the propagation launches each tile diagonal
    over the list of its tiles (pf_stage2_propagate_listed), or with
    dag=True runs the live tiles in one launch, each after the tiles it
    reads (pf_stage2_propagate_dag); then the sync, scan and stitch phases.
    Returns the device buffers (keep them alive until the stream has run)."""
    torch = _torch()
    lst, h_diag = diagonal_lists(plan, live, cache)
    dev = ws["tiles"].device
    d_list = torch.from_numpy(lst).to(dev)
    if dag:
        d_diag = torch.from_numpy(np.ascontiguousarray(h_diag, np.int64)).to(dev)
        work = torch.empty(plan.ntiles + 1, dtype=torch.int32, device=dev)
        _propagate_dag(plan, ws, d_list, d_diag, work, stream)
        stage2_phases(plan, sw, ws, _lib.PF_S2_SYNC | _lib.PF_S2_SCAN | _lib.PF_S2_STITCH, stream)
        return d_list, d_diag, work
    check(lib().pf_stage2_propagate_listed(
        ws["edge_C"].data_ptr(), ws["pairs"].data_ptr(), plan.npairs, ws["tiles"].data_ptr(),
        plan.ntiles, plan.L, d_list.data_ptr(), h_diag.ctypes.data, plan.max_diag,
        ws["edge_D"].data_ptr(), ws["edge_P"].data_ptr(), ws["D_arr"].data_ptr(),
        ws["L_arr"].data_ptr(), ws["olap"].data_ptr(), ws["bits"].data_ptr(), _stream(stream)),
        "pf_stage2_propagate_listed")
    stage2_phases(plan, sw, ws, _lib.PF_S2_SYNC | _lib.PF_S2_SCAN | _lib.PF_S2_STITCH, stream)
    return d_list


def _propagate_dag(plan, ws, d_list, d_diag, work, stream):
    check(lib().pf_stage2_propagate_dag(
        ws["edge_C"].data_ptr(), ws["pairs"].data_ptr(), plan.npairs, ws["tiles"].data_ptr(),
        plan.ntiles, plan.L, d_list.data_ptr(), d_diag.data_ptr(), plan.max_diag,
        ws["edge_D"].data_ptr(), ws["edge_P"].data_ptr(), ws["D_arr"].data_ptr(),
        ws["L_arr"].data_ptr(), ws["bits"].data_ptr(), work.data_ptr(), _stream(stream)),
        "pf_stage2_propagate_dag")


def stage2_listed_device(plan, sw, ws, d_list, h_diag, stream=None, d_diag=None, work=None):
    """stage2_listed with the diagonal lists already on the device (d_list,
    int32, one slot per tile) and their layout on the host (h_diag int64
    (max_diag, 3): first slot, live count, other count), as pf_select_tiles
    writes them.  With the layout on the device too (d_diag) and a work
    buffer of ntiles + 1 int32, the live tiles run as one tile-DAG launch
    (pf_stage2_propagate_dag)."""
    if d_diag is not None:
        _propagate_dag(plan, ws, d_list, d_diag, work, stream)
        stage2_phases(plan, sw, ws, _lib.PF_S2_SYNC | _lib.PF_S2_SCAN | _lib.PF_S2_STITCH, stream)
        return
    h_diag = np.ascontiguousarray(np.asarray(h_diag, np.int64).reshape(plan.max_diag, 3))
    check(lib().pf_stage2_propagate_listed(
        ws["edge_C"].data_ptr(), ws["pairs"].data_ptr(), plan.npairs, ws["tiles"].data_ptr(),
        plan.ntiles, plan.L, d_list.data_ptr(), h_diag.ctypes.data, plan.max_diag,
        ws["edge_D"].data_ptr(), ws["edge_P"].data_ptr(), ws["D_arr"].data_ptr(),
        ws["L_arr"].data_ptr(), ws["olap"].data_ptr(), ws["bits"].data_ptr(), _stream(stream)),
        "pf_stage2_propagate_listed")
    stage2_phases(plan, sw, ws, _lib.PF_S2_SYNC | _lib.PF_S2_SCAN | _lib.PF_S2_STITCH, stream)


def stitch_requests(plan, sw, ws, requests, stream=None):
    """backtrace_and_stitch (Parflex.py:1032-1103) from arbitrary edge slots.
    requests: int64 (n, 5) = (pair, ci, cj, et, idx).  Returns list of (n_i, 2)."""
    torch = _torch()
    req = np.ascontiguousarray(np.asarray(requests, dtype=np.int64).reshape(-1, 5))
    n = req.shape[0]
    if n == 0:
        return []
    dev = ws["pairs"].device
    d_req = torch.from_numpy(req).to(dev)
    cap = int(max(plan.pairs[req[:, 0], 0] + plan.pairs[req[:, 0], 1])) + 8
    chain = torch.empty(n * MAX_STITCH_TILES * 8, dtype=torch.int64, device=dev)
    pts = torch.empty(n * MAX_STITCH_TILES * (2 * plan.L + 8) * 2, dtype=torch.int64, device=dev)
    path = torch.empty(n * cap * 2, dtype=torch.int64, device=dev)
    plen = torch.empty(n, dtype=torch.int64, device=dev)
    rc = lib().pf_stitch(d_req.data_ptr(), n, ws["pairs"].data_ptr(), ws["tiles"].data_ptr(),
                         plan.L, sw.sp, sw.K, ws["Bp"].data_ptr(), ws["edge_P"].data_ptr(),
                         chain.data_ptr(), pts.data_ptr(), cap, path.data_ptr(),
                         plen.data_ptr(), _stream(stream))
    check(rc, "pf_stitch")
    path_h = path.cpu().numpy().reshape(n, cap, 2)
    plen_h = plen.cpu().numpy()
    return [path_h[q, :plen_h[q]].copy() for q in range(n)]


@dataclass
class BatchResult:
    best_cost: np.ndarray     # (P,) float64; NaN where no endpoint was found
    found: np.ndarray         # (P,) bool
    paths: list               # P arrays (2, n) int64 (wp layout of Parflex.parflex)
    best_end: np.ndarray      # (P, 6) (g_row, g_col, ci, cj, et, idx)


def collect(plan, ws):
    """Copy only best costs, endpoints and stitched paths back to the host
    (ws entries may also be host tensors already holding them)."""
    if ws["best_f"].is_cuda:
        device_status()
    best_f = ws["best_f"].cpu().numpy()
    best_i = ws["best_i"].cpu().numpy().reshape(plan.npairs, 8)
    plen = ws["path_len"].cpu().numpy()
    path = ws["path"].cpu().numpy().reshape(-1, 2)
    found = best_i[:, 6] == 1
    paths = []
    for p in range(plan.npairs):
        o = int(plan.pairs[p, 11])
        pts = path[o:o + int(plen[p])]
        paths.append(np.ascontiguousarray(pts.T))
    bc = np.where(found, best_f, np.nan)
    return BatchResult(bc, found, paths, best_i[:, :6].copy())


def collect_pinned(plan, ws, host):
    """collect() through pinned host buffers kept in the dict `host` (reused
    while the sizes match)."""
    torch = _torch()
    for k in ("best_f", "best_i", "path_len", "path"):
        if k not in host or host[k].shape != ws[k].shape:
            host[k] = torch.empty(ws[k].shape, dtype=ws[k].dtype, pin_memory=True)
        host[k].copy_(ws[k], non_blocking=True)
    torch.cuda.synchronize()
    device_status()
    return collect(plan, host)


def collect_packed(plan, ws, host, stream=None):
    """collect() with the paths packed on the device (pf_pack_paths): the
    small per-pair results go through pinned buffers kept in `host`, then one
    copy brings the 2 * sum(len) path values into a fresh pinned buffer, and
    every returned path is a (2, n) int64 view of it (no host-side copy)."""
    torch = _torch()
    st = stream or torch.cuda.current_stream()
    n = plan.npairs
    if "packed" not in ws:
        ws["packed"] = torch.empty(ws["path"].numel(), dtype=torch.int64,
                                   device=ws["path"].device)
    max_path = int((plan.pairs[:, 0] + plan.pairs[:, 1]).max()) + 8
    check(lib().pf_pack_paths(ws["pairs"].data_ptr(), n, max_path, ws["path"].data_ptr(),
                              ws["path_len"].data_ptr(), ws["packed"].data_ptr(), st.cuda_stream),
          "pf_pack_paths")
    with torch.cuda.stream(st):
        for k in ("best_f", "best_i", "path_len"):
            if k not in host or host[k].shape != ws[k].shape:
                host[k] = torch.empty(ws[k].shape, dtype=ws[k].dtype, pin_memory=True)
            host[k].copy_(ws[k], non_blocking=True)
    device_status(st)
    plen = host["path_len"].numpy().copy()
    total = 2 * int(plen.sum())
    buf = torch.empty(max(total, 1), dtype=torch.int64, pin_memory=True)
    if total:
        with torch.cuda.stream(st):
            buf[:total].copy_(ws["packed"][:total], non_blocking=True)
        st.synchronize()
    arr = buf.numpy()
    best_f = host["best_f"].numpy().copy()
    best_i = host["best_i"].numpy().reshape(n, 8).copy()
    found = best_i[:, 6] == 1
    paths, o = [], 0
    for p in range(n):
        m = int(plen[p])
        paths.append(arr[o:o + 2 * m].reshape(2, m))
        o += 2 * m
    bc = np.where(found, best_f, np.nan)
    return BatchResult(bc, found, paths, best_i[:, :6].copy())


def run_batch(C_dev, plan, steps, weights, stream=None, ws=None):
    """Stage 1 + stage 2 for every pair of the plan; returns (BatchResult, ws)."""
    sw = steps if isinstance(steps, StepWeights) else StepWeights(steps, weights)
    if ws is None:
        ws = plan.allocate(C_dev.device)
    stage1(C_dev, plan, sw, ws, stream)
    stage2(C_dev, plan, sw, ws, stream)
    return collect(plan, ws), ws


def single_pair_plan(L1, L2, L, beta):
    return BatchPlan([(int(L1), int(L2))], int(L), beta=float(beta))


# --------------------------------------------------------------------------
# batched pipeline: features -> cost -> stage 1 -> stage 2 for many pairs
# --------------------------------------------------------------------------

_TUNED = {}  # (L, dtype code, ntiles bucket) -> stage-1 launch shape


def _config_candidates(L, feat=False, Kp=16):
    """(strips, waves) shapes with 2 columns per lane and no wave idle (the
    features form keeps its C blocks in LDS: 8 KB per wave, at most 160 KB),
    and for the features form (K <= 128) one strip of 4 columns per lane
    (k_sweep4, tiles up to 4096 columns)."""
    out = []
    for nw in (1, 2, 4, 8, 16):
        if (nw - 1) * 128 < L and (not feat or nw * 8 <= 160):
            out.append((-(-L // (nw * 128)), nw))
    nw4 = -(-L // 256)
    if feat and Kp in (16, 32, 64, 128) and nw4 <= 16 and nw4 * 128 < L and (1, nw4) not in out:
        out.append((1, nw4))
    return out


def autotune_stage1(C_dev, plan, sw, ws, Gs=(32, 64), reps=2, stream=None, feat=None):
    """Time stage 1 for the candidate launch shapes on this plan and keep the
    fastest in ws['s1_config'] (and a process-wide cache keyed by tile length,
    dtype and tile count).  The exchange block G only matters with more than
    one strip; one-strip shapes are timed with the first G.  Returns
    (config, {config: ms})."""
    torch = _torch()
    key = (plan.L, "feat" if feat is not None else dtype_code(C_dev), plan.ntiles // 64)
    if key in _TUNED:
        ws["s1_config"] = _TUNED[key]
        return _TUNED[key], {}
    times = {}
    for (S, nw) in _config_candidates(plan.L, feat is not None,
                                      int(feat[2]) if feat is not None else 16):
        for G in (Gs if S > 1 else Gs[:1]):
            cfg = (S, nw, G)
            stage1(C_dev, plan, sw, ws, stream, config=cfg, feat=feat)  # warm-up / allocation
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                stage1(C_dev, plan, sw, ws, stream, config=cfg, feat=feat)
            e1.record()
            torch.cuda.synchronize()
            times[cfg] = e0.elapsed_time(e1) / reps
    best = min(times, key=times.get)
    _TUNED[key] = best
    ws["s1_config"] = best
    return best, times


class BatchRunner:
    """All pairs of a batch in one pipeline: cost matrices of every pair in one
    device buffer (one pf_cost per pair), one stage-1 launch over every tile
    of every pair, one stage-2 pass over the pair table.  The whole sequence
    can be captured in a HIP graph (capture()) and replayed (replay()): every
    buffer is allocated up front and every fill is a kernel."""

    def __init__(self, pairs, L, steps, weights, beta=0.1, dtype=None, device=None, share=None,
                 fused=False, s2_dag=False):
        """pairs: list of (F1 (K, N), F2 (K, M)) host feature arrays.  dtype of
        the cost matrices: torch.bfloat16 (default; FP32 DP) or float32 or
        float64 (exact plain form).  share: another runner over the same
        pairs and dtype whose features and cost buffer are reused (a sweep
        over tile lengths builds the cost matrices once).  fused=True (BF16
        only): no cost matrix is stored; stage 1 computes C from the BF16
        features inside the sweep (pf_stage1_features).  s2_dag: stage 2's
        propagation as one tile-DAG launch instead of a launch per tile
        diagonal (the same results; with every tile live the per-diagonal
        launches are faster: 256 pairs at L = 1000, 2.2 against 2.6 ms)."""
        torch = _torch()
        self.s2_dag = bool(s2_dag)
        self.sw = StepWeights(steps, weights)
        shapes = [(int(F1.shape[1]), int(F2.shape[1])) for F1, F2 in pairs]
        # Rows of every cost matrix start on a 16-byte boundary (row stride a
        # multiple of 8 elements) so the cost kernel writes whole 16-byte chunks.
        ldcs = [-(-M // 8) * 8 for _, M in shapes]
        offs = np.concatenate([[0], np.cumsum([N * ld for (N, _), ld in zip(shapes, ldcs)])])
        self.plan = BatchPlan(shapes, int(L), beta=float(beta),
                              c_offsets=[int(o) for o in offs[:-1]], ldcs=ldcs)
        if share is not None:
            if [tuple(x) for x in share.plan.shapes] != shapes:
                raise ValueError("share: runner over different pairs")
            self.dtype, self.device = share.dtype, share.device
            self.C, self.F1, self.F2, self.K = share.C, share.F1, share.F2, share.K
            self.work, self.code, self.jobs = share.work, share.code, share.jobs
            self.cost_blocks = share.cost_blocks
            self.c_ready = share.c_ready
            self.fused, self.feat = share.fused, share.feat
        else:
            self.dtype = torch.bfloat16 if dtype is None else dtype
            self.device = device or torch.device("cuda", torch.cuda.current_device())
            self.fused = bool(fused)
            if self.fused and self.dtype != torch.bfloat16:
                raise ValueError("the fused form computes the cost in BF16")
            # the fused form stores no cost matrix
            self.C = torch.empty(1 if self.fused else int(offs[-1]), dtype=self.dtype,
                                 device=self.device)
            self.code = dtype_code(self.C)
            K = int(pairs[0][0].shape[0])
            if any(int(F1.shape[0]) != K or int(F2.shape[0]) != K for F1, F2 in pairs):
                raise ValueError("all feature arrays of a batch need the same feature count K")
            # features of every pair concatenated along frames: (K, sum N), (K, sum M)
            self.F1 = torch.from_numpy(np.ascontiguousarray(
                np.concatenate([F1 for F1, _ in pairs], axis=1), dtype=np.float64)).to(self.device)
            self.F2 = torch.from_numpy(np.ascontiguousarray(
                np.concatenate([F2 for _, F2 in pairs], axis=1), dtype=np.float64)).to(self.device)
            self.K = K
            tr, tc = ctypes.c_int64(), ctypes.c_int64()
            check(lib().pf_cost_tile(self.code, ctypes.byref(tr), ctypes.byref(tc)), "pf_cost_tile")
            jobs = np.zeros((len(pairs), 8), np.int64)
            f1 = f2 = blk = 0
            for p, (N, M) in enumerate(shapes):
                jobs[p] = [f1, f2, N, M, self.plan.pairs[p, 5], self.plan.pairs[p, 6], blk, 0]
                f1 += N
                f2 += M
                blk += (-(-N // tr.value)) * (-(-M // tc.value))
            self.cost_blocks = blk
            self.jobs = torch.from_numpy(jobs).to(self.device)
            nb = ctypes.c_int64()
            check(lib().pf_cost_workspace(self.code, K, f1, f2, ctypes.byref(nb)),
                  "pf_cost_workspace")
            self.work = torch.empty(-(-nb.value // 8), dtype=torch.float64, device=self.device)
            self.c_ready = [False]
            self.feat = None
            if self.fused:
                Kp = -(-K // 16) * 16
                self.feat = (torch.empty(f1 * Kp, dtype=torch.bfloat16, device=self.device),
                             torch.empty(f2 * Kp, dtype=torch.bfloat16, device=self.device), Kp)
        self.ws = self.plan.allocate(self.device)
        if self.fused:
            # tile table columns 0/1: frames of the tile's first row / column
            t = self.plan.tiles.copy()
            f1o = np.concatenate([[0], np.cumsum([a for a, _ in shapes])[:-1]])
            f2o = np.concatenate([[0], np.cumsum([b for _, b in shapes])[:-1]])
            hop = int(L) - 1
            t[:, 0] = f1o[t[:, 6]] + (t[:, 7] >> 32) * hop
            t[:, 1] = f2o[t[:, 6]] + (t[:, 7] & 0xFFFFFFFF) * hop
            self.ws["tiles"] = torch.from_numpy(t).to(self.device)
        self.graph = None

    def cost(self, stream=None):
        """Cost matrices of every pair: one batched launch (pf_cost_batch);
        in the fused form, the BF16 normalised features instead."""
        if self.fused:
            F1b, F2b, _ = self.feat
            for F, Fb in ((self.F1, F1b), (self.F2, F2b)):
                check(lib().pf_features_bf16(F.data_ptr(), self.K, int(F.shape[1]), Fb.data_ptr(),
                                             _stream(stream)), "pf_features_bf16")
            self.c_ready[0] = True
            return
        check(lib().pf_cost_batch(self.code, self.F1.data_ptr(), int(self.F1.shape[1]),
                                  self.F2.data_ptr(), int(self.F2.shape[1]), self.K,
                                  self.jobs.data_ptr(), int(self.jobs.shape[0]), self.cost_blocks,
                                  self.C.data_ptr(), self.work.data_ptr(), _stream(stream)),
              "pf_cost_batch")
        self.c_ready[0] = True

    def launch(self, stream=None):
        """Cost build, stage 1 and stage 2 for every pair (asynchronous)."""
        self.cost(stream)
        stage1(self.C, self.plan, self.sw, self.ws, stream, feat=self.feat)
        stage2(self.C, self.plan, self.sw, self.ws, stream, dag=self.s2_dag)

    def autotune(self, Gs=(32,)):
        if not self.c_ready[0]:
            self.cost()
        return autotune_stage1(self.C, self.plan, self.sw, self.ws, Gs=Gs, feat=self.feat)

    def time_launch(self, reps=2):
        """Mean ms of stage 1 + stage 2 (the cost build does not depend on L)."""
        torch = _torch()
        if not self.c_ready[0]:
            self.cost()
        stage1(self.C, self.plan, self.sw, self.ws, feat=self.feat)
        stage2(self.C, self.plan, self.sw, self.ws, dag=self.s2_dag)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            stage1(self.C, self.plan, self.sw, self.ws, feat=self.feat)
            stage2(self.C, self.plan, self.sw, self.ws, dag=self.s2_dag)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    def capture(self):
        """Record the three phases (cost build, stage 1, stage 2) as three HIP
        graphs, after a warm-up run that sizes the strip-exchange buffers."""
        torch = _torch()
        self.launch()
        torch.cuda.synchronize()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        graphs = []
        with torch.cuda.stream(s):
            for phase in (self.cost,
                          lambda st: stage1(self.C, self.plan, self.sw, self.ws, st, feat=self.feat),
                          lambda st: stage2(self.C, self.plan, self.sw, self.ws, st,
                                            dag=self.s2_dag)):
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=s):
                    phase(s)
                graphs.append(g)
        torch.cuda.current_stream().wait_stream(s)
        self.graph = graphs
        return graphs

    def replay(self, events=None):
        """Replay the captured phases on the current stream; with events (a
        list of 4 torch.cuda.Event) record one before and after each phase."""
        torch = _torch()
        st = torch.cuda.current_stream()
        if events is not None:
            events[0].record(st)
        for k, g in enumerate(self.graph):
            g.replay()
            if events is not None:
                events[k + 1].record(st)

    def collect(self):
        """Best costs, endpoints and stitched paths of every pair, copied
        through pinned host buffers allocated on first use."""
        if getattr(self, "_host", None) is None:
            self._host = {}
        return collect_packed(self.plan, self.ws, self._host)


def tune_batch(pairs, Ls, steps, weights, beta=0.1, dtype=None, Gs=(32,), reps=2, log=None,
               fused=False):
    """Autotuner over the tile length L and the stage-1 launch shape
    (strips, waves, exchange block G) for one batch: for every L the fastest
    launch shape is found (autotune_stage1), then the L whose stage 1 +
    stage 2 is fastest is kept.  The cost matrices are built once and shared.
    Returns (runner, {L: (config, ms)})."""
    best, best_ms, report, base = None, None, {}, None
    for L in Ls:
        br = BatchRunner(pairs, L, steps, weights, beta, dtype, share=base, fused=fused)
        if base is None:
            base = br
        cfg, _ = br.autotune(Gs=Gs)
        ms = br.time_launch(reps)
        report[int(L)] = (cfg, ms)
        if log is not None:
            log(f"L={L}: tiles={br.plan.ntiles} config={cfg} stage1+stage2 {ms:.2f} ms")
        if best is None or ms < best_ms:
            best, best_ms = br, ms
        elif br is not base:
            del br
    return best, report
