"""Batched ParFlex with learned tile proposal and full-tile fallback.

For every pair the tile network (tile_net.TileNet) scores the tiles from the
1/32-pooled cost; the tiles whose logit reaches the threshold, dilated by a
margin, form the selection.  Then:

  * the cost is built only for selected tiles, one slot per tile
    (pf_cost_batch with one job per tile, BF16);
  * stage 1 runs on the selected tiles only (a work table: the selected rows
    of the full tile table, pointing at their cost slots);
  * stage 2 runs on the full tile grid of every pair; pruned tiles have
    D = +inf on their edges, so no candidate or chain passes through them.
    Their back-pointer words point at one shared all-zero region, so any
    read of them stays in bounds.

Selection: a tile is kept when its score -- the largest logit over its pooled
cells extended by `ext` cells on every side -- reaches the threshold, so the
margin around the predicted path is ext*32 frames, not whole tiles.

A pair falls back to the plain full-tile method (engine.BatchRunner over all
its tiles) when a tile within `guard` frames of its stitched path was pruned
(the path reached the edge of the selection), when no endpoint is found, or
when its normalised cost exceeds the calibrated bound.  The plain method is
also the reference the result is checked against.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from . import engine
from ._lib import PF_BF16, check, lib
from ._plan import BatchPlan
from .engine import StepWeights, _stream
from .tile_net import POOL, TileNet

__all__ = ["LearnedRunner", "path_tiles", "needs_fallback"]


def _torch():
    import torch
    return torch


def path_tiles(wp, L, n1, n2, guard=0, L1=None, L2=None):
    """Flat indices (i*n2 + j, unique) of the tiles holding a cell within
    `guard` frames (in row and column) of a point of the path (2, n); a cell
    on a tile overlap belongs to every tile that holds it."""
    hop = L - 1
    r0 = np.asarray(wp[0], np.int64)
    c0 = np.asarray(wp[1], np.int64)
    if r0.size == 0:
        return np.zeros(0, np.int64)
    offs = (0,) if guard <= 0 else (-guard, 0, guard)
    rmax = (L1 - 1) if L1 is not None else np.iinfo(np.int64).max
    cmax = (L2 - 1) if L2 is not None else np.iinfo(np.int64).max

    def tiles_of(v, vmax, n):
        out = []
        for d in offs:
            x = np.clip(v + d, 0, vmax)
            a = np.minimum(x // hop, n - 1)
            b = np.minimum(np.where((x % hop == 0) & (x > 0), x // hop - 1, x // hop), n - 1)
            out.extend((a, b))
        return out

    rows = tiles_of(r0, rmax, n1)
    cols = tiles_of(c0, cmax, n2)
    flat = np.concatenate([ri * n2 + cj for ri in rows for cj in cols])
    return np.unique(flat)


def needs_fallback(wp, found, best_cost, sel, L, cost_bound, guard, L1, L2):
    """Reason the learned result cannot be trusted, or None: no endpoint,
    normalised cost above the bound, or a tile within `guard` frames of the
    path was pruned (the path reached the edge of the selection)."""
    if not found or not np.isfinite(best_cost):
        return "no endpoint"
    if best_cost > cost_bound:
        return "cost bound"
    if _pruned_near_path(wp, sel, L, guard, L1, L2):
        return "selection edge"
    return None


def _pruned_near_path(wp, sel, L, guard, L1, L2):
    """True when a pruned tile holds a cell within `guard` frames of a path
    point.  The tiles holding frames [x0, x1] are max(x0-1, 0)//hop ..
    min(x1//hop, n-1) (a frame on an overlap belongs to two tiles), so for
    each point they form a rectangle of tile indices, tested with a prefix
    sum of the pruned mask."""
    if np.asarray(wp).size == 0:
        return False
    # Every step moves at most 2 frames in each coordinate, so testing every
    # 8th point (and the last) with the guard widened by 8 frames covers all
    # points with the original guard (a conservative test).
    n = np.asarray(wp).shape[1]
    idx = np.unique(np.concatenate([np.arange(0, n, 8), [n - 1]]))
    r = np.asarray(wp[0])[idx].astype(np.int32)
    c = np.asarray(wp[1])[idx].astype(np.int32)
    n1, n2 = sel.shape
    hop = np.int32(L - 1)
    g = np.int32(guard + 8)
    i0 = np.minimum(np.maximum(r - g - 1, 0) // hop, n1 - 1)
    i1 = np.minimum(np.minimum(r + g, L1 - 1) // hop, n1 - 1)
    j0 = np.minimum(np.maximum(c - g - 1, 0) // hop, n2 - 1)
    j1 = np.minimum(np.minimum(c + g, L2 - 1) // hop, n2 - 1)
    pr = np.zeros((n1 + 1, n2 + 1), np.int32)
    pr[1:, 1:] = np.cumsum(np.cumsum(~sel, axis=0), axis=1)
    w = n2 + 1
    flat = pr.reshape(-1)
    cnt = (flat[(i1 + 1) * w + j1 + 1] - flat[i0 * w + j1 + 1] - flat[(i1 + 1) * w + j0]
           + flat[i0 * w + j0])
    return bool(cnt.any())


class LearnedRunner:
    """One batch of feature pairs at tile length L through the learned
    pipeline.  run() does one complete pass (scoring, selection, pruned
    stage 1 + stage 2, fallback) and returns (BatchResult, stats)."""

    def __init__(self, pairs, L, steps, weights, beta, net: TileNet, threshold, ext=8,
                 guard=128, cost_bound=math.inf, device=None, graph=True, fused=True,
                 fused_net=True):
        torch = _torch()
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.L, self.beta = int(L), float(beta)
        self.steps, self.weights = steps, weights
        self.sw = StepWeights(steps, weights)
        self.net, self.threshold = net, float(threshold)
        self.ext, self.guard = int(ext), int(guard)
        if not 0 <= 2 * self.guard < self.L - 1:
            raise ValueError("guard must be below half the tile hop (L-1)/2")
        self.cost_bound = float(cost_bound)
        self.pairs = pairs
        self.shapes = [(int(F1.shape[1]), int(F2.shape[1])) for F1, F2 in pairs]
        K = int(pairs[0][0].shape[0])
        if K > 16:
            raise ValueError("the learned form takes at most 16 features per frame")
        self.K = K
        f64 = np.float64
        self.Fp = [(torch.from_numpy(np.ascontiguousarray(F1, f64)).to(self.device),
                    torch.from_numpy(np.ascontiguousarray(F2, f64)).to(self.device))
                   for F1, F2 in pairs]
        self.F1 = torch.cat([a for a, _ in self.Fp], dim=1).contiguous()
        self.F2 = torch.cat([b for _, b in self.Fp], dim=1).contiguous()
        self.f1_off = np.concatenate([[0], np.cumsum([s[0] for s in self.shapes])[:-1]])
        self.f2_off = np.concatenate([[0], np.cumsum([s[1] for s in self.shapes])[:-1]])
        nb = ctypes.c_int64()
        check(lib().pf_cost_workspace(PF_BF16, K, int(self.F1.shape[1]), int(self.F2.shape[1]),
                                      ctypes.byref(nb)), "pf_cost_workspace")
        self.work = torch.empty(-(-nb.value // 8), dtype=torch.float64, device=self.device)
        self.plan = BatchPlan(self.shapes, self.L, beta=self.beta)
        tr, tc = ctypes.c_int64(), ctypes.c_int64()
        check(lib().pf_cost_tile(PF_BF16, ctypes.byref(tr), ctypes.byref(tc)), "pf_cost_tile")
        self.cost_tile = (tr.value, tc.value)
        self._host = {}
        self.graph, self._graph, self._scores = bool(graph), None, None
        self.nstreams, self._side, self._bufs = 2, None, None
        # pooled blocks of every pair, concatenated: pair p's rows are columns
        # [hb[p], hb[p] + H_p) of the K x sum(H) array, its columns likewise
        self.HW = [(-(-a // POOL), -(-b // POOL)) for a, b in self.shapes]
        H_, W_ = np.array([h for h, _ in self.HW]), np.array([w for _, w in self.HW])
        self.hb = np.concatenate([[0], np.cumsum(H_)[:-1]]).astype(np.int64)
        self.wb = np.concatenate([[0], np.cumsum(W_)[:-1]]).astype(np.int64)
        self.nbA, self.nbB = int(H_.sum()), int(W_.sum())
        n = len(pairs)
        seg1 = np.zeros((n, 4), np.int64)
        seg1[:, 0], seg1[:, 1], seg1[:, 2] = self.f1_off, [a for a, _ in self.shapes], self.hb
        seg2 = np.zeros((n, 4), np.int64)
        seg2[:, 0], seg2[:, 1], seg2[:, 2] = self.f2_off, [b for _, b in self.shapes], self.wb
        stp = np.stack([self.hb, H_, self.wb, W_], axis=1).astype(np.int64)
        self._seg = [torch.from_numpy(t).to(self.device) for t in (seg1, seg2, stp)]
        self.Ap = torch.empty(K * self.nbA, dtype=torch.float32, device=self.device)
        self.Bp = torch.empty(K * self.nbB, dtype=torch.float32, device=self.device)
        self.stats = torch.empty(n * (2 * K + 1), dtype=torch.float32, device=self.device)
        # fused_net: the four layers of every pair in one launch (pf_tile_net)
        # and the tile maxima of every pair in one more (pf_tile_scores_batch);
        # otherwise one launch per layer and pair (the per-layer kernels, which
        # give the same logits bit for bit)
        self.fused_net = bool(fused_net)
        if self.fused_net:
            self._net_tables()
        # fused: the selected tiles' cost is computed inside stage 1 from the
        # BF16 features (pf_stage1_features) instead of per-tile cost slots
        self.fused = bool(fused)
        if self.fused:
            Kp = -(-K // 16) * 16
            self.feat = (torch.empty(int(self.F1.shape[1]) * Kp, dtype=torch.bfloat16,
                                     device=self.device),
                         torch.empty(int(self.F2.shape[1]) * Kp, dtype=torch.bfloat16,
                                     device=self.device), Kp)

    def _net_tables(self):
        """Pair and job tables of pf_tile_net and pf_tile_scores_batch (the
        layouts of include/parflex_hip.h) and the logit maps of every pair."""
        torch = _torch()
        cj, rj = ctypes.c_int64(), ctypes.c_int64()
        check(lib().pf_tile_net_geometry(ctypes.byref(cj), ctypes.byref(rj)),
              "pf_tile_net_geometry")
        n = len(self.pairs)
        K = self.K
        npair = np.zeros((n, 8), np.int64)
        spair = np.zeros((n, 8), np.int64)
        jobs, tpair = [], []
        z_off = 0
        for p, ((H, W), (L1, L2)) in enumerate(zip(self.HW, self.shapes)):
            npair[p] = [self.hb[p], H, self.wb[p], W, z_off, p * (2 * K + 1), 0, 0]
            n1, n2 = self.plan.grids[p]
            spair[p] = [H, W, z_off, L1, L2, n2, self.plan.pairs[p, 4], 0]
            for x0 in range(0, W, cj.value):
                for y0 in range(0, H, rj.value):
                    jobs.append((p, x0, y0, min(y0 + rj.value, H)))
            tpair.append(np.full(n1 * n2, p, np.int32))
            z_off += H * W
        dev = self.device
        self._npairs_t = torch.from_numpy(npair).to(dev)
        self._spairs_t = torch.from_numpy(spair).to(dev)
        self._jobs_t = torch.from_numpy(np.array(jobs, np.int64).reshape(-1, 4)).to(dev)
        self._njobs = len(jobs)
        self._tpair_t = torch.from_numpy(np.concatenate(tpair)).to(dev)
        self._Z = torch.empty(max(z_off, 1), dtype=torch.float32, device=dev)

    def logit_maps(self, stream=None):
        """Logit map (H, W) of every pair as host arrays (scoring pass
        without the tile maxima), from the fused kernel or the per-layer
        kernels as the runner is built."""
        torch = _torch()
        out = torch.empty(int(self.plan.ntiles), dtype=torch.float32, device=self.device)
        maps = self._score_into(out, stream, keep_maps=True)
        torch.cuda.synchronize()
        return [m.cpu().numpy().reshape(h, w) for m, (h, w) in zip(maps, self.HW)]

    # ------------------------------------------------------------------
    def _score_into(self, out, stream=None, keep_maps=False):
        """Tile logits of every pair into the device buffer `out` (the
        pairs' grids concatenated in i-major order).  Pooled features (in
        the fused form with the BF16 features stage 1 reads, from the same
        pass) and input statistics of all pairs take three launches on
        `stream`; the
        network layers of the pairs are spread over side streams (forked
        from and joined back to `stream`) so that one pair's layers overlap
        the tail of another's."""
        torch = _torch()
        main = stream or torch.cuda.current_stream()
        K, n = self.K, len(self.pairs)
        seg1, seg2, stp = self._seg
        ms = main.cuda_stream
        if self.fused:
            # the fused form's BF16 features come from the same pass over F
            F1b, F2b, _ = self.feat
            for F, sg, P, nbp, Fb in ((self.F1, seg1, self.Ap, self.nbA, F1b),
                                      (self.F2, seg2, self.Bp, self.nbB, F2b)):
                check(lib().pf_pool_features_bf16_batch(F.data_ptr(), K, int(F.shape[1]),
                                                        sg.data_ptr(), n, POOL, P.data_ptr(),
                                                        nbp, Fb.data_ptr(), ms),
                      "pf_pool_features_bf16_batch")
        else:
            check(lib().pf_pool_features_batch(self.F1.data_ptr(), K, int(self.F1.shape[1]),
                                               seg1.data_ptr(), n, POOL, self.Ap.data_ptr(),
                                               self.nbA, ms), "pf_pool_features_batch")
            check(lib().pf_pool_features_batch(self.F2.data_ptr(), K, int(self.F2.shape[1]),
                                               seg2.data_ptr(), n, POOL, self.Bp.data_ptr(),
                                               self.nbB, ms), "pf_pool_features_batch")
        check(lib().pf_input_stats_batch(self.Ap.data_ptr(), self.nbA, self.Bp.data_ptr(),
                                         self.nbB, K, stp.data_ptr(), n, self.stats.data_ptr(),
                                         ms), "pf_input_stats_batch")
        if self.fused_net:
            net = self.net
            bias = [net.layer(i)[32 * 288:] for i in range(4)]
            check(lib().pf_tile_net(self.Ap.data_ptr(), self.nbA, self.Bp.data_ptr(), self.nbB, K,
                                    self.stats.data_ptr(), self._npairs_t.data_ptr(), n,
                                    self._jobs_t.data_ptr(), self._njobs,
                                    net.Wt[0].data_ptr(), bias[0].data_ptr(),
                                    net.Wt[1].data_ptr(), bias[1].data_ptr(),
                                    net.Wt[2].data_ptr(), bias[2].data_ptr(),
                                    net.Wt[3].data_ptr(), bias[3].data_ptr(),
                                    self._Z.data_ptr(), ms), "pf_tile_net")
            check(lib().pf_tile_scores_batch(self._Z.data_ptr(), self._spairs_t.data_ptr(),
                                             self._tpair_t.data_ptr(), int(self.plan.ntiles),
                                             POOL, self.L, self.ext, out.data_ptr(), ms),
                  "pf_tile_scores_batch")
            if keep_maps:
                zo = self._spairs_t.cpu().numpy()[:, 2]
                return [self._Z[int(o):int(o) + h * w] for o, (h, w) in zip(zo, self.HW)]
            return None
        ns = self.nstreams
        if self._side is None:
            self._side = [torch.cuda.Stream() for _ in range(ns)]
        if self._bufs is None:
            # activation buffers per side stream, sized for its largest pair
            self._bufs = []
            for q in range(ns):
                hw = max([h * w for p, (h, w) in enumerate(self.HW) if p % ns == q] or [1])
                self._bufs.append((torch.empty(hw * 32, dtype=torch.bfloat16, device=self.device),
                                   torch.empty(hw * 32, dtype=torch.bfloat16, device=self.device),
                                   torch.empty(hw, dtype=torch.float32, device=self.device)))
        for sd in self._side:
            sd.wait_stream(main)
        st0 = self.stats.data_ptr()
        o = 0
        maps = []
        for p in range(n):
            sd = self._side[p % ns]
            H, W = self.HW[p]
            L1, L2 = self.shapes[p]
            if keep_maps:
                bufs = (self._bufs[p % ns][0], self._bufs[p % ns][1],
                        torch.empty(H * W, dtype=torch.float32, device=self.device))
            else:
                bufs = self._bufs[p % ns]
            Z = self.net.forward_pooled(self.Ap, self.nbA, int(self.hb[p]), H, self.Bp, self.nbB,
                                        int(self.wb[p]), W, K, st0 + 4 * p * (2 * K + 1),
                                        bufs, sd)
            if keep_maps:
                maps.append(Z)
            n1, n2 = -(-(L1 - 1) // (self.L - 1)), -(-(L2 - 1) // (self.L - 1))
            check(lib().pf_tile_scores(Z.data_ptr(), H, W, POOL, self.L, L1, L2, self.ext,
                                       out[o:o + n1 * n2].data_ptr(), sd.cuda_stream),
                  "pf_tile_scores")
            o += n1 * n2
        for sd in self._side:
            main.wait_stream(sd)
        return maps if keep_maps else None

    def score(self, stream=None):
        """Tile logits of every pair, as one host array per pair (n1, n2).
        With graph=True the scoring of all pairs is recorded once as a HIP
        graph (after one eager pass) and replayed; the network's weights are
        then fixed for this runner."""
        flat = self._score_flat(stream)
        res, o = [], 0
        for (n1, n2) in self.plan.grids:
            res.append(flat[o:o + n1 * n2].reshape(n1, n2))
            o += n1 * n2
        return res

    def _score_flat(self, stream=None):
        """Tile logits of every pair as one host array, the pairs' grids
        concatenated in i-major order (the plan's tile order)."""
        return self._score_dev(stream).cpu().numpy()

    def _score_dev(self, stream=None):
        """Tile logits of every pair on the device (the tensor _score_flat
        copies to the host)."""
        torch = _torch()
        ntot = int(self.plan.ntiles)
        if self._scores is None:
            self._scores = torch.empty(ntot, dtype=torch.float32, device=self.device)
        if not self.graph:
            self._score_into(self._scores, stream)
        else:
            if self._graph is None:
                self._score_into(self._scores, stream)
                torch.cuda.synchronize()
                g = torch.cuda.CUDAGraph()
                s = torch.cuda.Stream()
                s.wait_stream(torch.cuda.current_stream())
                with torch.cuda.stream(s):
                    with torch.cuda.graph(g, stream=s):
                        self._score_into(self._scores, s)
                torch.cuda.current_stream().wait_stream(s)
                self._graph = g
            self._graph.replay()
        return self._scores

    def _sel_tables(self):
        """Device tables and buffers of pf_select_tiles, made once: the plan's
        tile and pair tables, the pairs' first feature frames, the first list
        slot of every tile diagonal, and the outputs (selection, tile tables,
        diagonal lists, counts with their pinned host copy)."""
        if getattr(self, "_sel", None) is not None:
            return self._sel
        torch = _torch()
        plan, dev = self.plan, self.device
        ij = plan.tiles[:, 7]
        dg = (ij >> 32) + (ij & 0xFFFFFFFF)
        per = np.bincount(dg, minlength=plan.max_diag)
        diag_off = np.concatenate([[0], np.cumsum(per)]).astype(np.int64)
        frames = np.stack([self.f1_off, self.f2_off], axis=1).astype(np.int64)
        n = int(plan.ntiles)
        ncount = 3 + 3 * plan.max_diag
        self._sel = {
            "plan": torch.from_numpy(plan.tiles).to(dev),
            "pairs": torch.from_numpy(plan.pairs).to(dev),
            "frames": torch.from_numpy(np.ascontiguousarray(frames)).to(dev),
            "diag_off": torch.from_numpy(diag_off).to(dev),
            "sel": torch.empty(n, dtype=torch.uint8, device=dev),
            "full": torch.empty((n, 8), dtype=torch.int64, device=dev),
            "work": torch.empty((n, 8), dtype=torch.int64, device=dev),
            "list": torch.empty(n, dtype=torch.int32, device=dev),
            "counts": torch.empty(ncount, dtype=torch.int64, device=dev),
            "tmp": torch.empty(3 * (-(-n // 1024)) + 2 * plan.max_diag, dtype=torch.int64,
                               device=dev),
            "h_counts": torch.empty(ncount, dtype=torch.int64, pin_memory=True),
            # done flags and ticket counter of the stage-2 tile DAG
            "dag": torch.empty(n + 1, dtype=torch.int32, device=dev),
        }
        return self._sel

    def _select_device(self, d_scores, stream=None):
        """select_tiles on the device (pf_select_tiles) for the fused form:
        the tables stay on the device and only the counts come back.
        Returns (nsel, b_words, cells, h_diag)."""
        torch = _torch()
        t = self._sel_tables()
        plan = self.plan
        sobj = stream or torch.cuda.current_stream()
        check(lib().pf_select_tiles(d_scores.data_ptr(), self.threshold, t["plan"].data_ptr(),
                                    plan.ntiles, t["frames"].data_ptr(), self.L,
                                    t["diag_off"].data_ptr(), plan.max_diag, t["sel"].data_ptr(),
                                    t["full"].data_ptr(), t["work"].data_ptr(),
                                    t["list"].data_ptr(), t["counts"].data_ptr(),
                                    t["tmp"].data_ptr(), sobj.cuda_stream), "pf_select_tiles")
        with torch.cuda.stream(sobj):
            t["h_counts"].copy_(t["counts"], non_blocking=True)
        sobj.synchronize()
        c = t["h_counts"].numpy()
        return int(c[0]), int(c[1]), int(c[2]), c[3:].reshape(-1, 3).copy()

    def _pruned_tables(self, sels):
        """Work table, full table with remapped back-pointer offsets, cost
        jobs and sizes for the selection (one boolean per tile of the plan,
        flat in the plan's order, or one (n1, n2) array per pair)."""
        plan = self.plan
        full = plan.tiles.copy()
        sel = (np.asarray(sels, bool) if isinstance(sels, np.ndarray)
               else np.concatenate([np.asarray(s).reshape(-1) for s in sels]))
        R, W = full[:, 2], full[:, 3]
        words = ((R + 15) // 16) * W
        zero_words = int(((self.L + 15) // 16) * self.L)
        b_off = np.zeros(len(full), np.int64)
        ws = words[sel]
        b_off[sel] = zero_words + np.concatenate([[0], np.cumsum(ws)[:-1]]) if ws.size else []
        full[:, 4] = b_off                    # pruned tiles read the zero region
        full[~sel, 0] = -1                    # pruned: no stage 1, edges +inf (pf_stage2)
        work = full[sel].copy()
        ldc = (W[sel] + 7) // 8 * 8
        slot = R[sel] * ldc
        c_off = np.concatenate([[0], np.cumsum(slot)[:-1]]) if slot.size else np.zeros(0, np.int64)
        # the tile's frames in the concatenated feature arrays
        pair = work[:, 6]
        ij = work[:, 7]
        hop = self.L - 1
        s1 = (ij >> 32) * hop
        s2 = (ij & 0xFFFFFFFF) * hop
        if self.fused:
            work[:, 0] = self.f1_off[pair] + s1      # frame of the tile's first row
            work[:, 1] = self.f2_off[pair] + s2      # frame of the tile's first column
        else:
            work[:, 0] = c_off
            work[:, 1] = ldc
        tr, tc = self.cost_tile
        nblk = (-(-work[:, 2] // tr)) * (-(-work[:, 3] // tc))
        jobs = np.zeros((len(work), 8), np.int64)
        jobs[:, 0] = self.f1_off[pair] + s1
        jobs[:, 1] = self.f2_off[pair] + s2
        jobs[:, 2] = work[:, 2]
        jobs[:, 3] = work[:, 3]
        jobs[:, 4] = c_off
        jobs[:, 5] = ldc
        jobs[:, 6] = np.concatenate([[0], np.cumsum(nblk)[:-1]]) if len(work) else []
        return (sel, full, work, jobs, int(nblk.sum()), zero_words + int(ws.sum()),
                int(slot.sum()), int((R[sel] * W[sel]).sum()))

    def run(self, stream=None, timings=None, events=None):
        """One complete pass.  With a dict for `timings`, the device is
        synchronised after every phase and the phase times (s) are added.
        With a pair of torch.cuda.Event for `events`, they are recorded on the
        launch stream around the stage-1 kernel.  With a stream other than
        the current one, the whole pass (fills, uploads, fallback included)
        runs with that stream made current."""
        torch = _torch()
        if stream is not None and stream != torch.cuda.current_stream():
            with torch.cuda.stream(stream):
                return self.run(None, timings, events)
        st = _stream(stream)
        L = self.L
        clock = [time.perf_counter()]

        def mark(name):
            if timings is not None:
                torch.cuda.synchronize()
                now = time.perf_counter()
                timings[name] = timings.get(name, 0.0) + now - clock[0]
                clock[0] = now

        dev = self.device
        plan = self.plan
        if self.fused:
            # select_tiles with no margin on every pair at once (score >=
            # threshold), on the device: tables, lists and selection stay there
            d_scores = self._score_dev(stream)
            mark("score")
            nsel, b_words, cells, h_diag = self._select_device(d_scores, stream)
            t = self._sel_tables()
            mark("select")
            ws = plan.allocate(dev, b_words=b_words, tiles=t["full"], pairs=t["pairs"])
        else:
            flat = self._score_flat(stream)
            mark("score")
            sel, full, work, jobs, nblk, b_words, c_elems, cells = self._pruned_tables(
                flat >= self.threshold)
            mark("select")
            nsel = len(work)
            ws = plan.allocate(dev, b_words=b_words, tiles=torch.from_numpy(full).to(dev))
        # Bp holds the zero region (read by pruned tiles) + the selected tiles;
        # edge arrays need no fill: stage 2 does not read pruned tiles' slots
        ws["Bp"][:(L + 15) // 16 * L].zero_()
        mark("workspace")
        if nsel:
            if self.fused:
                # the BF16 features were written by the scoring pass's pooling
                C = None
                F1b, F2b, Kp = self.feat
            else:
                C = torch.empty(c_elems, dtype=torch.bfloat16, device=dev)
                d_jobs = torch.from_numpy(jobs).to(dev)
                check(lib().pf_cost_batch(PF_BF16, self.F1.data_ptr(), int(self.F1.shape[1]),
                                          self.F2.data_ptr(), int(self.F2.shape[1]), self.K,
                                          d_jobs.data_ptr(), nsel, nblk, C.data_ptr(),
                                          self.work.data_ptr(), st), "pf_cost_batch")
            mark("cost")
            d_work = self._sel["work"] if self.fused else torch.from_numpy(work).to(dev)
            S, nw, G = self._s1_shape(nsel)
            fl = xc = None
            if S > 1:
                nf, nx = ctypes.c_int64(), ctypes.c_int64()
                check(lib().pf_stage1_workspace(PF_BF16, nsel, L, S, ctypes.byref(nf),
                                                ctypes.byref(nx)), "pf_stage1_workspace")
                flags = torch.empty(nf.value, dtype=torch.int32, device=dev)
                xch = torch.empty(-(-nx.value // 8), dtype=torch.float64, device=dev)
                fl, xc = flags.data_ptr(), xch.data_ptr()
            if events is not None:
                events[0].record(stream or torch.cuda.current_stream())
            if self.fused:
                check(lib().pf_stage1_features(F1b.data_ptr(), F2b.data_ptr(), Kp,
                                               d_work.data_ptr(), nsel, L, self.sw.sp, self.sw.wp,
                                               self.sw.K, S, nw, G, fl, xc, ws["Bp"].data_ptr(),
                                               ws["edge_D"].data_ptr(), ws["edge_P"].data_ptr(),
                                               ws["edge_C"].data_ptr(), st), "pf_stage1_features")
            else:
                check(lib().pf_stage1(PF_BF16, C.data_ptr(), d_work.data_ptr(), nsel, L,
                                      self.sw.sp, self.sw.wp, self.sw.K, S, nw, G, fl, xc,
                                      ws["Bp"].data_ptr(), ws["edge_D"].data_ptr(),
                                      ws["edge_P"].data_ptr(), ws["edge_C"].data_ptr(), st),
                      "pf_stage1")
            if events is not None:
                events[1].record(stream or torch.cuda.current_stream())
            mark("stage1")
        if self.fused:
            t = self._sel
            engine.stage2_listed_device(plan, self.sw, ws, t["list"], h_diag, stream,
                                        d_diag=t["counts"][3:], work=t["dag"])
            d_sel = self._sel["sel"]
        else:
            self._stage2_listed(sel, ws, st)
            d_sel = torch.from_numpy(sel.astype(np.uint8)).to(dev)
        mark("stage2")
        # selection check on the device: a pruned tile near a path
        flag = torch.zeros(len(self.pairs), dtype=torch.int32, device=dev)
        max_path = max(a + b for a, b in self.shapes) + 8
        check(lib().pf_selection_check(ws["pairs"].data_ptr(), len(self.pairs), max_path,
                                       ws["path"].data_ptr(), ws["path_len"].data_ptr(),
                                       d_sel.data_ptr(), L, self.guard, flag.data_ptr(), st),
              "pf_selection_check")
        res = engine.collect_packed(plan, ws, self._host, stream)
        edge = flag.cpu().numpy()
        mark("collect")
        fb, why = [], {}
        for p in range(len(self.pairs)):
            bc = float(res.best_cost[p])
            reason = ("no endpoint" if not res.found[p] or not np.isfinite(bc) else
                      "cost bound" if bc > self.cost_bound else
                      "selection edge" if edge[p] else None)
            if reason:
                fb.append(p)
                why[reason] = why.get(reason, 0) + 1
        mark("check")
        if fb:
            sub = [self.pairs[p] for p in fb]
            br = engine.BatchRunner(sub, L, self.steps, self.weights, self.beta,
                                    torch.bfloat16, dev)
            br.launch(stream)
            r2 = br.collect()
            for q, p in enumerate(fb):
                res.best_cost[p] = r2.best_cost[q]
                res.found[p] = r2.found[q]
                res.paths[p] = r2.paths[q]
                res.best_end[p] = r2.best_end[q]
            mark("fallback")
        stats = {"tiles": int(plan.ntiles), "selected_tiles": nsel,
                 "cells_full": int(plan.cells), "cells_computed": cells,
                 "fallback_pairs": fb, "fallback_reasons": why}
        return res, stats

    def _stage2_listed(self, sel, ws, st):
        """Stage 2 with the tile diagonals launched over lists of their
        tiles, selected tiles first; the same results as engine.stage2."""
        if getattr(self, "_diag_cache", None) is None:
            self._diag_cache = {}
        self._d_list = engine.stage2_listed(self.plan, self.sw, ws, sel, cache=self._diag_cache)

    def _s1_shape(self, ntiles):
        """Stage-1 launch shape (strips, waves, G) for a large selection: one
        strip of 4 columns per lane per tile up to 4096 columns (the fused
        form then runs k_sweep4: at L = 2000, 8 waves, 14.9 ms against 22.0
        ms for one strip of 16 waves with 2 columns per lane), else 2 columns
        per lane in strips of up to 16 waves with 64-row exchange blocks."""
        L = self.L
        if ntiles >= 512 and self.fused and L <= 4096:
            return 1, -(-L // 256), 64
        if ntiles >= 512:
            strips = -(-L // 2048)
            nw = max(1, min(16, -(-L // (128 * strips))))
            return -(-L // (nw * 128)), nw, 64
        return engine.stage1_config(self.plan)
