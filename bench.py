"""ParFlex benchmark on one MI355X.

Workloads (--workload):

  chroma_learned128_N20000-60000_L2000 (default): 128 held-out seeded chroma
      pairs (seeds 90000+p, N ~ U[20000, 60000], M = N * U[0.8, 1.25],
      trimmed ends), tile length 2000.  The learned form: the tile network
      (parflexdtw_amd/data/tile_net_L2000.npz, trained on 10k other pairs)
      scores every tile from the 1/32-pooled cost, only the selected tiles
      run stage 1 (FP32 DP, the BF16 MFMA cost computed inside the sweep
      from the features; --cost materialised builds it first), stage 2
      (FP64) runs on the full grid, and a pair whose path reaches the edge of its selection
      (or exceeds the calibrated cost bound) is redone by the full-tile
      batched form.  One step = features (in HBM) -> paths on the host for
      all pairs.  Beside it: the plain FP64 drop-in (parflex per pair, cost
      build included) on every pair, the reference for agreement.
  chroma_batch256_N3000-20000: 256 seeded chroma pairs (K = 12),
      N ~ U[3000, 20000], M = N * U[0.8, 1.25], trimmed ends, so every pair
      has ragged edge tiles.  The batched form: one stage-1 launch over every
      tile of every pair (FP32 DP, the BF16 MFMA cost computed inside the
      sweep; --cost materialised builds all 256 cost matrices first), one
      stage-2 pass (FP64 scan), captured as HIP graphs;
      the tuner picks L in {500, 1000, 2000, 4000} and the stage-1 launch
      shape.  One step = features (resident in HBM) -> cost -> stage 1 ->
      stage 2 -> best costs and paths on the host for all 256 pairs.
      Beside it: the plain FP64 drop-in (parflex per pair, cost build
      included) on a sample of the same pairs at the same L, whose paths are
      the reference for agreement (frames within +-2), and the CPU oracle on the
      same sample.
  chroma16000x19000_L4000: the plain FP64 drop-in on one pair, exact against
      the CPU oracle (cost matrix resident in HBM).
  relu128_40000x40000_L4000: the precision form on one dense pair (K = 128):
      BF16 MFMA cost + FP32 DP + FP64 stage 2, with cost error and path
      agreement against the FP64 plain form.
  chroma2000x2400_L500: the small plain case.

Printed JSON (one line) follows the contract of the repository; the roofline
figure is for the stage-1 DP kernel (the dominant kernel), algorithmic bytes
per cell = one read of C + 2 bits of packed back-pointer, timed with HIP
events on the stream it runs on.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DP GCUPS + pairs/s, batched/learned vs plain GPU at >=99.5% path agreement"
# HBM bytes per stage-1 launch measured with rocprofv3 PMC (FETCH_SIZE +
# WRITE_SIZE, scaled by a 1 GiB copy); profiles/pmc_traffic_*.md hold the runs.
TRAFFIC_BYTES = {("chroma_batch256_N3000-20000", 500, "materialised"): (83.97 + 12.84) * 1e9,
                 ("chroma_learned128_N20000-60000_L2000", 2000, "materialised"):
                     (46.70 + 6.21) * 1e9,
                 ("chroma_learned128_N20000-60000_L2000", 2000, "fused"): (0.743 + 6.148) * 1e9}
# VALU wave instructions per computed cell of the stage-1 kernel, measured
# with rocprofv3 PMC (SQ_INSTS_VALU / cells; profiles/pmc_valu_*.md).  The
# bench states the kernel's VALU instruction rate per SIMD and cycle (at the
# 2.4 GHz engine clock, the clock the PMC cycle count shows for this kernel)
# beside the rate a dense FP32 FMA loop (pf_valu_probe) reaches in the same
# run.
VALU_INSTS_PER_CELL = {("chroma_learned128_N20000-60000_L2000", 2000, "fused"):
                           8.8873e9 / 2.2747e10}
HBM_PEAK_GBS = 8000.0
STEPS, WEIGHTS, BETA = [(1, 1), (1, 2), (2, 1)], [2, 3, 3], 0.1
WORKLOADS = ["chroma_learned128_N20000-60000_L2000", "chroma_batch256_N3000-20000",
             "chroma16000x19000_L4000", "relu128_40000x40000_L4000", "chroma2000x2400_L500"]
MODEL = os.path.join(ROOT, "parflexdtw_amd", "data", "tile_net_L2000.npz")


def _parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default=WORKLOADS[0], choices=WORKLOADS)
    ap.add_argument("--pairs", type=int, default=None,
                    help="batch size (default 128 learned, 256 batched)")
    ap.add_argument("--L", type=int, nargs="*", default=[500, 1000, 2000, 4000],
                    help="tile lengths the batched workload's tuner tries")
    ap.add_argument("--plain-sample", type=int, default=16,
                    help="pairs of the batch also run through the plain FP64 form")
    ap.add_argument("--cost", choices=["fused", "materialised", "auto"], default=None,
                    help="precision forms: compute C inside stage 1 from BF16 features (fused), "
                         "build the BF16 cost matrix first (materialised), or time both and "
                         "keep the faster (auto; the default of the single-pair precision "
                         "workload, fused for the others)")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-trials", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one 1 GiB device-to-device copy, the known-size reference "
                         "for scaling HBM counters (FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.cost is None:
        args.cost = "auto" if args.workload == "relu128_40000x40000_L4000" else "fused"
    elif args.cost == "auto" and args.workload != "relu128_40000x40000_L4000":
        ap.error("--cost auto applies to the single-pair precision workload only")
    return args


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _cpu_model():
    try:
        return [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo")
                if l.startswith("model name")][0]
    except Exception:
        return "unknown"


def _cpu_timed(oracle, Cs, L, nthreads, trials):
    """runtime.py:144-173's method: perf_counter around each whole parflex
    call, one warm-up pass over the sample (dropped, as the reference drops
    its first trials, runtime_analysis.ipynb#cell2), then `trials` timed
    passes.  Returns (seconds per timed pass, results of the last pass)."""
    times, res = [], []
    for k in range(trials + 1):
        res = []
        t0 = time.perf_counter()
        for C in Cs:
            res.append(oracle.parflex(C, STEPS, WEIGHTS, BETA, L, nthreads=nthreads))
        if k > 0:
            times.append(time.perf_counter() - t0)
    return times, res


def _cpu_baseline(args, Cs, Cs1, L, what, per_unit, unit, expect=None):
    """The CPU baseline (the FP64 oracle, a port of the reference's
    parflex): the cpu_parflex.py:392 schedule (stage 1 on a pool of
    min(tiles, threads) workers, stage 2 on one thread) on the sample Cs, and
    the single-thread form (Parflex.py:18-22 with one thread, runtime.py:144)
    on the sample Cs1.  `per_unit` converts cells/s into the metric's unit
    (pairs/s at the workload's mean pair size, or GCUPS).  `expect` holds the
    plain GPU results of Cs, which the CPU results must equal."""
    from oracle import oracle
    cores = len(os.sched_getaffinity(0))
    nthreads = max(1, min(args.cpu_threads, cores))
    trials = max(1, args.cpu_trials)
    cells = sum(C.shape[0] * C.shape[1] for C in Cs)
    cells1 = sum(C.shape[0] * C.shape[1] for C in Cs1)
    t_m, res = _cpu_timed(oracle, Cs, L, nthreads, trials)
    t_1, _ = _cpu_timed(oracle, Cs1, L, 1, trials)
    med_m, med_1 = float(np.median(t_m)), float(np.median(t_1))
    exact = None
    if expect is not None:
        exact = all(bo == be and np.array_equal(wpo, wpe)
                    for (bo, wpo), (be, wpe) in zip(res, expect))
    out = {
        "value": round(per_unit(cells / med_m), 4), "unit": unit, "cores": nthreads,
        "kind": "port",
        "sample": f"{what}: {len(Cs)} pair(s), {cells / 1e9:.3f} Gcells, FP64 parflex end to "
                  f"end at L={L} on {nthreads} worker threads (cpu_parflex.py:392 schedule), "
                  f"median of {trials} timed passes after 1 warm-up pass",
        "seconds": round(med_m, 4), "trials_s": [round(t, 4) for t in t_m],
        "gcups": round(cells / med_m / 1e9, 4),
        "affinity_cores": cores, "cpu_model": _cpu_model(),
        "single_thread": {
            "value": round(per_unit(cells1 / med_1), 4), "unit": unit, "cores": 1,
            "gcups": round(cells1 / med_1 / 1e9, 4), "seconds": round(med_1, 4),
            "trials_s": [round(t, 4) for t in t_1],
            "sample": f"{len(Cs1)} pair(s), {cells1 / 1e9:.3f} Gcells, one thread, median of "
                      f"{trials} timed passes after 1 warm-up pass"}}
    return out, exact


def _feat_kernel(cfg, L, K=12):
    """Name of the features-form stage-1 kernel a launch shape runs."""
    S, nw = int(cfg[0]), int(cfg[1])
    cpl = 2 if S * nw * 128 >= L else 4
    if S == 1 and cpl == 4 and K <= 128:
        return f"k_sweep4<1024,{-(-K // 16)}>"
    return f"k_stage1<{cpl},__hip_bfloat16,float,false,true,true,{1024 if cpl == 2 else 512}>"


def _valu_probe_rate(torch):
    """VALU wave instructions per second of pf_valu_probe: 8 workgroups of
    4 waves per CU (256 CUs), each wave 16 FMAs x iters, timed with HIP
    events on the launch stream (best of 3)."""
    from parflexdtw_amd._lib import check, lib
    blocks, iters = 2048, 20000
    out = torch.empty(blocks * 256, dtype=torch.float32, device="cuda")
    st = torch.cuda.current_stream()
    best = None
    for _ in range(4):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        check(lib().pf_valu_probe(out.data_ptr(), blocks, iters, st.cuda_stream), "pf_valu_probe")
        e1.record(st)
        st.synchronize()
        ms = e0.elapsed_time(e1)
        best = ms if best is None else min(best, ms)
    return blocks * 4 * 16 * iters / (best * 1e-3)


def _roofline(cells, s1_ms, bytes_per_cell, kernel):
    achieved = cells * bytes_per_cell / (s1_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "kernel": kernel,
            "bytes_per_cell": bytes_per_cell}


def _base(args, value, unit, ms_step, dtype, config):
    return {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": None,
            "config": config}


# ---------------------------------------------------------------------------
EVAL_TOL_MS = (25, 50, 100, 200, 500)


def _evaluation(torch, gts, hyps, runner=None):
    """Error rate versus tolerance (05_Evaluate_Benchmarks_L.ipynb#cell23)
    of each set of paths in `hyps` against the synthetic ground-truth beats,
    computed by the device evaluator; with `runner`, also the evaluator read
    in place from the runner's device paths, its launch timed with HIP
    events (median of 5)."""
    from parflexdtw_amd.evaluate import HOP_SEC, evaluate_batch, evaluate_paths
    out = {"metric": "error rate at tolerance (fraction of beats with |error| > t), "
                     "05_Evaluate_Benchmarks_L.ipynb#cell23",
           "ground_truth": "synthetic beats every 22 F2 frames (0.51 s) with their F1 "
                           "position from the known warp; hop 512/22050 s"}
    for name, (paths, g) in hyps.items():
        r = evaluate_paths(paths, g, HOP_SEC, 1000)
        out[name] = {"beats": r.beats,
                     **{f"{t}ms": round(float(r.rates[t]), 6) for t in EVAL_TOL_MS}}
    if runner is not None:
        st = torch.cuda.current_stream()
        ts = []
        for _ in range(5):
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            r = evaluate_batch(runner, gts, HOP_SEC, 1000, stream=st, events=ev)
            ts.append(ev[0].elapsed_time(ev[1]))
        us = float(np.median(ts)) * 1e3
        out["device_paths"] = {"beats": r.beats,
                               **{f"{t}ms": round(float(r.rates[t]), 6) for t in EVAL_TOL_MS}}
        out["kernel"] = {"name": "k_eval_paths", "us": round(us, 2),
                         "beats_per_s": round(r.beats / (us * 1e-6), 1),
                         "note": "paths read in place from stage 2's device buffer; one "
                                 "bisection per beat, 1001 thresholds"}
    return out


def bench_learned(args, torch):
    from parflexdtw_amd import Parflex_gpu as PG
    from parflexdtw_amd.gpu_flexdtw import cost_matrix_from_features_gpu
    from parflexdtw_amd.learned import LearnedRunner
    from parflexdtw_amd.evaluate import HOP_SEC
    from parflexdtw_amd.synth import batch_pairs, beat_times, path_agreement
    from parflexdtw_amd.tile_net import TileNet

    P = args.pairs or 128
    net, meta = TileNet.load(MODEL)
    L = int(meta["L"])
    t0 = time.perf_counter()
    trip = batch_pairs(P, 90000, 20000, 60000)
    pairs = [(a, b) for a, b, _ in trip]
    gts = [beat_times(w, HOP_SEC, every=22) for _, _, w in trip]
    del trip
    shapes = [(F1.shape[1], F2.shape[1]) for F1, F2 in pairs]
    log(f"{P} held-out pairs (seeds 90000+, N ~ U[20000,60000]) in "
        f"{time.perf_counter() - t0:.1f} s; {sum(a * b for a, b in shapes) / 1e9:.1f} Gcells; "
        f"network trained on {int(meta['train_pairs'])} pairs, threshold "
        f"{float(meta['threshold']):.3f}")
    lr = LearnedRunner(pairs, L, STEPS, WEIGHTS, BETA, net, float(meta["threshold"]),
                       int(meta["ext"]), int(meta["guard"]), float(meta["cost_bound"]),
                       fused=args.cost == "fused")
    for _ in range(args.warmup):
        res, st = lr.run()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        res, st = lr.run(events=evs[k])
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / args.steps
    s1_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in evs])) if st["selected_tiles"] else 0.0
    tm = {}
    lr.run(timings=tm)
    pairs_s = P / (ms_step * 1e-3)
    log(f"learned: {ms_step:.2f} ms/step ({pairs_s:.1f} pairs/s); phases (synchronised run): "
        + ", ".join(f"{k} {v * 1e3:.1f}" for k, v in tm.items()))
    log(f"selected {st['selected_tiles']}/{st['tiles']} tiles, "
        f"{st['cells_computed'] / st['cells_full']:.4f} of the cells; fallbacks "
        f"{len(st['fallback_pairs'])} {st['fallback_reasons']}")

    # the plain FP64 drop-in on every pair: the reference for agreement, and the plain pairs/s
    F1d, F2d = pairs[0]
    PG.parflex(cost_matrix_from_features_gpu(F1d, F2d), STEPS, WEIGHTS, BETA, L,
               profile_dir=None)
    torch.cuda.synchronize()
    plain = []
    t0 = time.perf_counter()
    for F1, F2 in pairs:
        C64 = cost_matrix_from_features_gpu(F1, F2)
        plain.append(PG.parflex(C64, STEPS, WEIGHTS, BETA, L, profile_dir=None))
        del C64
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t0
    plain_pairs_s = P / plain_s
    agree = [path_agreement(plain[p][1], res.paths[p], tol=2) for p in range(P)]
    rel = [abs(res.best_cost[p] - plain[p][0]) / abs(plain[p][0]) for p in range(P)]
    frames = sum(plain[p][1].shape[1] for p in range(P))
    frames_ok = sum(path_agreement(plain[p][1], res.paths[p], tol=2) * plain[p][1].shape[1]
                    for p in range(P))
    log(f"plain FP64 drop-in: {plain_s:.2f} s for {P} pairs ({plain_pairs_s:.2f} pairs/s); "
        f"agreement min {min(agree):.5f}, all frames {frames_ok / frames:.6f}")
    cells = st["cells_full"]
    out = _base(args, round(pairs_s, 3), "pairs/s", ms_step, "bf16 cost, f32 DP, f64 stage 2",
                {"workload": args.workload, "pairs": P, "N": "U[20000,60000]",
                 "M": "N*U[0.8,1.25], 0-10% trimmed each end", "K": 12, "L": L,
                 "tiles": st["tiles"], "cells": cells, "steps": STEPS, "weights": WEIGHTS,
                 "beta": BETA, "form": "learned tile proposal + batched BF16/FP32 pipeline "
                                       f"({args.cost} cost), fallback to the plain method"})
    out["data"] = ("synthetic (seeded chroma pairs with known time warps, held out from the "
                   "network's training seeds; features in HBM)")
    out["speedup_vs_plain"] = round(pairs_s / plain_pairs_s, 3)
    out["plain_gpu"] = {"pairs_per_s": round(plain_pairs_s, 3),
                        "gcups": round(cells / plain_s / 1e9, 3), "pairs": P,
                        "form": "parflex drop-in per pair, FP64 cost + FP64 DP, same L"}
    out["effective_gcups"] = round(cells / (ms_step * 1e-3) / 1e9, 3)
    out["dp_gcups"] = round(st["cells_computed"] / (s1_ms * 1e-3) / 1e9, 3) if s1_ms else None
    out["stage1_ms"] = round(s1_ms, 4)
    out["phases_ms"] = {k: round(v * 1e3, 3) for k, v in tm.items()}
    out["selection"] = {"tiles_selected": st["selected_tiles"], "tiles": st["tiles"],
                        "cells_fraction": round(st["cells_computed"] / cells, 5),
                        "fallback_pairs": len(st["fallback_pairs"]),
                        "fallback_reasons": st["fallback_reasons"]}
    out["agreement"] = {"frames_within_2_all": round(frames_ok / frames, 6),
                        "frames_within_2_min_pair": round(min(agree), 6),
                        "best_cost_rel_err_max": float(max(rel)), "pairs_compared": P,
                        "reference": "plain FP64 path of the same pair at the same L"}
    out["model"] = {"file": "parflexdtw_amd/data/tile_net_L2000.npz",
                    "train_pairs": int(meta["train_pairs"]),
                    "train_N": f"U[{int(meta['n_lo'])},{int(meta['n_hi'])}]",
                    "epochs": int(meta["epochs_done"]), "threshold": float(meta["threshold"]),
                    "ext_cells": int(meta["ext"]), "guard_frames": int(meta["guard"])}
    out["evaluation"] = _evaluation(torch, gts, {
        "learned": (res.paths, gts), "plain_fp64": ([wp for _, wp in plain], gts)})
    fused = args.cost == "fused"
    kname = ("k_sweep4<1024,1>" if fused else
             "k_stage1<2,__hip_bfloat16,float,false,true,false,1024>")
    out["roofline"] = _roofline(st["cells_computed"], s1_ms, 0.25 if fused else 2.25,
                                kname) if s1_ms else None
    key = (args.workload, L, args.cost)
    if out["roofline"] and P == 128 and key in TRAFFIC_BYTES:
        out["roofline"]["traffic"] = TRAFFIC_BYTES[key]
    if out["roofline"] and P == 128 and key in VALU_INSTS_PER_CELL:
        ipc = VALU_INSTS_PER_CELL[key]
        rate = ipc * st["cells_computed"] / (s1_ms * 1e-3)
        probe = _valu_probe_rate(torch)
        per_cyc = 1024 * 2.4e9
        out["roofline"]["valu_issue"] = {
            "insts_per_cell": round(ipc, 4), "insts_per_s": round(rate, -6),
            "per_simd_cycle": round(rate / per_cyc, 4),
            "fma_probe_per_simd_cycle": round(probe / per_cyc, 4),
            "note": "VALU wave instructions of the stage-1 kernel (per cell from a PMC run of "
                    "this workload) per SIMD and 2.4 GHz cycle, and those of a dense FP32 FMA "
                    "loop (pf_valu_probe) in the same run; the DP issues more than the FMA loop "
                    "(which runs at a power-limited clock), so it is bound by its row "
                    "dependency chain rather than by issue alone"}
    if not args.no_cpu:
        # The three smallest of the first 8 held-out pairs (about 2 Gcells;
        # multi-thread) and the smallest of them (single thread), each checked
        # against the plain GPU path; one pair's C is up to ~7 GB on the host.
        first = list(range(min(8, P)))
        small = sorted(first, key=lambda p: shapes[p][0] * shapes[p][1])[:3]
        Cs = [cost_matrix_from_features_gpu(*pairs[p]).cpu().numpy() for p in small]
        mean_cells = sum(a * b for a, b in shapes) / P
        out["cpu_baseline"], exact = _cpu_baseline(
            args, Cs, Cs[:1], L, f"held-out pairs {small}", lambda cps: cps / mean_cells,
            "pairs/s", expect=[plain[p] for p in small])
        out["cpu_baseline"]["sample"] += (f"; pairs/s = cells/s / {mean_cells / 1e9:.3f} Gcells, "
                                          "the mean pair of the workload")
        del Cs
        out["parity"] = {"plain_gpu_vs_cpu_oracle_exact": bool(exact), "pairs": small}
    return out


# ---------------------------------------------------------------------------
def bench_batch(args, torch):
    from parflexdtw_amd import Parflex_gpu as PG
    from parflexdtw_amd import engine
    from parflexdtw_amd.gpu_flexdtw import cost_matrix_from_features_gpu
    from parflexdtw_amd.evaluate import HOP_SEC
    from parflexdtw_amd.synth import batch_pairs, beat_times, path_agreement

    t0 = time.perf_counter()
    args.pairs = args.pairs or 256
    trip = batch_pairs(args.pairs)
    pairs = [(F1, F2) for F1, F2, _ in trip]
    gts = [beat_times(w, HOP_SEC, every=22) for _, _, w in trip]
    del trip
    shapes = [(F1.shape[1], F2.shape[1]) for F1, F2 in pairs]
    log(f"{len(pairs)} pairs generated in {time.perf_counter() - t0:.1f} s; "
        f"{sum(a * b for a, b in shapes) / 1e9:.2f} Gcells")
    t0 = time.perf_counter()
    br, report = engine.tune_batch(pairs, args.L, STEPS, WEIGHTS, BETA, torch.bfloat16, log=log,
                                   fused=args.cost == "fused")
    L = br.plan.L
    log(f"tuned in {time.perf_counter() - t0:.1f} s: L={L} config={br.ws['s1_config']}")
    br.capture()
    for _ in range(args.warmup):
        br.replay()
        res = br.collect()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        br.replay(evs[k])
        res = br.collect()
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / args.steps
    cost_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
    s1_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))
    s2_ms = float(np.mean([e[2].elapsed_time(e[3]) for e in evs]))
    cells = br.plan.cells
    P = len(pairs)
    pairs_s = P / (ms_step * 1e-3)
    log(f"batched: {ms_step:.2f} ms/step ({pairs_s:.1f} pairs/s), cost {cost_ms:.2f} ms, "
        f"stage1 {s1_ms:.2f} ms, stage2 {s2_ms:.2f} ms, found {int(res.found.sum())}/{P}")

    # Plain FP64 drop-in on a sample of the same pairs, same L: the reference for
    # agreement and the plain pairs/s.
    ns = min(args.plain_sample, P)
    plain, agree, rel_err = [], [], []
    Cs_host = []
    for p in range(ns):
        F1, F2 = pairs[p]
        C64 = cost_matrix_from_features_gpu(F1, F2)
        PG.parflex(C64, STEPS, WEIGHTS, BETA, L, profile_dir=None)  # warm-up for this shape
        del C64
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for p in range(ns):
        F1, F2 = pairs[p]
        C64 = cost_matrix_from_features_gpu(F1, F2)
        bc, wp = PG.parflex(C64, STEPS, WEIGHTS, BETA, L, profile_dir=None)
        plain.append((bc, wp))
        del C64
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t0
    if not args.no_cpu:
        Cs_host = [cost_matrix_from_features_gpu(*pairs[p]).cpu().numpy() for p in range(min(4, ns))]
    plain_pairs_s = ns / plain_s
    plain_cells = sum(shapes[p][0] * shapes[p][1] for p in range(ns))
    for p in range(ns):
        bc, wp = plain[p]
        agree.append(path_agreement(wp, res.paths[p], tol=2))
        rel_err.append(abs(res.best_cost[p] - bc) / max(abs(bc), 1e-300))
    log(f"plain FP64 on {ns} pairs: {plain_s * 1e3:.1f} ms ({plain_pairs_s:.1f} pairs/s); "
        f"agreement min {min(agree):.5f} mean {np.mean(agree):.5f}")

    out = _base(args, round(pairs_s, 3), "pairs/s", ms_step, "bf16 cost, f32 DP, f64 stage 2",
                {"workload": args.workload, "pairs": P, "N": "U[3000,20000]",
                 "M": "N*U[0.8,1.25], 0-10% trimmed each end", "K": 12, "L": L,
                 "L_tried": sorted(report), "stage1_shape": list(br.ws["s1_config"]),
                 "tiles": br.plan.ntiles, "cells": cells, "steps": STEPS, "weights": WEIGHTS,
                 "beta": BETA,
                 "form": ("batched: BF16 MFMA cost computed inside stage 1 (fused), FP32 DP, "
                          "FP64 stage 2, HIP graphs" if args.cost == "fused" else
                          "batched: BF16 MFMA cost matrix, FP32 DP, FP64 stage 2, HIP graphs")})
    out["data"] = "synthetic (seeded chroma pairs with known time warps, features in HBM)"
    out["dp_gcups"] = round(cells / (s1_ms * 1e-3) / 1e9, 3)
    out["gcups_end_to_end"] = round(cells / (ms_step * 1e-3) / 1e9, 3)
    out["cost_ms"] = round(cost_ms, 4)
    out["stage1_ms"] = round(s1_ms, 4)
    out["stage2_ms"] = round(s2_ms, 4)
    out["tuner"] = {str(k): {"config": list(v[0]), "stage1_stage2_ms": round(v[1], 3)}
                    for k, v in sorted(report.items())}
    out["plain_gpu"] = {"pairs_per_s": round(plain_pairs_s, 3),
                        "gcups": round(plain_cells / plain_s / 1e9, 3), "sample_pairs": ns,
                        "form": "parflex drop-in per pair, FP64 cost + FP64 DP, same L"}
    out["speedup_vs_plain"] = round(pairs_s / plain_pairs_s, 3)
    out["agreement"] = {"frames_within_2_min": round(min(agree), 6),
                        "frames_within_2_mean": round(float(np.mean(agree)), 6),
                        "best_cost_rel_err_max": float(max(rel_err)),
                        "pairs_compared": ns, "reference": "plain FP64 path, same pair and L"}
    fused = args.cost == "fused"
    out["roofline"] = _roofline(
        cells, s1_ms, 0.25 if fused else 2.25,
        _feat_kernel(br.ws["s1_config"], L) if fused else
        "k_stage1<2,__hip_bfloat16,float,false,true,false,1024>")
    if args.pairs == 256 and (args.workload, L, args.cost) in TRAFFIC_BYTES:
        out["roofline"]["traffic"] = TRAFFIC_BYTES[(args.workload, L, args.cost)]
    out["evaluation"] = _evaluation(torch, gts, {
        "batched_sample": (res.paths[:ns], gts[:ns]),
        "plain_fp64_sample": ([wp for _, wp in plain], gts[:ns])}, runner=br)

    if not args.no_cpu:
        mean_cells = sum(a * b for a, b in shapes) / P
        small = sorted(range(len(Cs_host)), key=lambda p: Cs_host[p].size)[:1]
        out["cpu_baseline"], exact = _cpu_baseline(
            args, Cs_host, [Cs_host[p] for p in small], L,
            f"pairs 0-{len(Cs_host) - 1} of the batch", lambda cps: cps / mean_cells, "pairs/s",
            expect=plain)
        out["cpu_baseline"]["sample"] += (f"; pairs/s = cells/s / {mean_cells / 1e9:.3f} Gcells, "
                                          "the mean pair of the workload")
        out["parity"] = {"plain_gpu_vs_cpu_oracle_exact": bool(exact), "pairs": len(Cs_host)}
    return out


# ---------------------------------------------------------------------------
def bench_plain(args, torch):
    from parflexdtw_amd import engine
    from parflexdtw_amd._plan import BatchPlan
    from parflexdtw_amd.gpu_flexdtw import cost_matrix_from_features_gpu
    from parflexdtw_amd.synth import cfg_pair, make_pair

    if args.workload == "chroma16000x19000_L4000":
        F1, F2, _ = cfg_pair(1)
        L = 4000
    else:
        F1, F2, _ = make_pair(2000, 2400, seed=0)
        L = 500
    dev = torch.device("cuda", 0)
    C_dev = cost_matrix_from_features_gpu(F1, F2)
    L1, L2 = int(C_dev.shape[0]), int(C_dev.shape[1])
    plan = BatchPlan([(L1, L2)], L, beta=BETA)
    sw = engine.StepWeights(STEPS, WEIGHTS)
    ws = plan.allocate(dev)
    stream = torch.cuda.current_stream()
    log(f"workload {args.workload}: {L1}x{L2}, L={L}, tiles={plan.ntiles}, cells={plan.cells}")

    def step(ev=None):
        if ev is not None:
            ev[0].record(stream)
        engine.stage1(C_dev, plan, sw, ws, stream)
        if ev is not None:
            ev[1].record(stream)
        engine.stage2(C_dev, plan, sw, ws, stream)
        if ev is not None:
            ev[2].record(stream)
        return engine.collect(plan, ws)

    for _ in range(args.warmup):
        res = step()
    torch.cuda.synchronize()
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        res = step(events[k])
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / args.steps
    s1_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in events]))
    s2_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in events]))
    cells = plan.cells
    log(f"step {ms_step:.3f} ms, stage1 {s1_ms:.3f} ms, stage2 {s2_ms:.3f} ms")
    out = _base(args, round(cells / (ms_step * 1e-3) / 1e9, 3), "GCUPS", ms_step, "f64",
                {"workload": args.workload, "L1": L1, "L2": L2, "L": L, "tiles": plan.ntiles,
                 "cells": cells, "steps": STEPS, "weights": WEIGHTS, "beta": BETA,
                 "form": "plain FP64 drop-in, stage1+stage2 on device"})
    out["data"] = "synthetic (seeded chroma pair with known time warp)"
    out["dp_gcups"] = round(cells / (s1_ms * 1e-3) / 1e9, 3)
    out["pairs_per_s"] = round(1e3 / ms_step, 3)
    out["stage1_ms"] = round(s1_ms, 4)
    out["stage2_ms"] = round(s2_ms, 4)
    out["roofline"] = _roofline(cells, s1_ms, 8.25, "k_stage1<2,double,double,false,true,false,256>")
    if not args.no_cpu:
        C = C_dev.cpu().numpy()
        out["cpu_baseline"], exact = _cpu_baseline(
            args, [C], [C], L, f"the whole {L1}x{L2} pair", lambda cps: cps / 1e9, "GCUPS",
            expect=[(float(res.best_cost[0]), res.paths[0])])
        out["parity"] = {"path_exact": bool(exact), "best_cost_gpu": float(res.best_cost[0]),
                         "path_len": int(res.paths[0].shape[1])}
    return out


# ---------------------------------------------------------------------------
def bench_precision(args, torch):
    from parflexdtw_amd import Parflex_gpu as PG
    from parflexdtw_amd import engine
    from parflexdtw_amd._plan import BatchPlan
    from parflexdtw_amd.gpu_flexdtw import cost_matrix_from_features_gpu
    from parflexdtw_amd.synth import cfg_pair, path_agreement

    F1, F2, _ = cfg_pair(3)
    L = 4000
    dev = torch.device("cuda", 0)
    F1d = torch.from_numpy(F1).to(dev)
    F2d = torch.from_numpy(F2).to(dev)
    L1, L2 = F1.shape[1], F2.shape[1]
    plan = BatchPlan([(L1, L2)], L, beta=BETA)
    log(f"workload {args.workload}: {L1}x{L2}, K={F1.shape[0]}, L={L}, tiles={plan.ntiles}")

    # Both forms run as three HIP graphs (cost, stage 1, stage 2) with the
    # stage-1 launch shape from the autotuner: fused computes C inside stage 1
    # from the BF16 features (no C in HBM), materialised builds the BF16 cost
    # matrix first (rows 16-byte aligned) and stage 1 reads it.  auto times one
    # pass of each and keeps the faster.
    forms = ["fused", "materialised"] if args.cost in (None, "auto") else [args.cost]
    runners, pass_ms = {}, {}
    for form in forms:
        r = engine.BatchRunner([(F1, F2)], L, STEPS, WEIGHTS, BETA, torch.bfloat16, dev,
                               fused=form == "fused")
        cfg_f, _ = r.autotune()
        r.capture()
        for _ in range(2):
            r.replay()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(3):
            r.replay()
        e1.record()
        torch.cuda.synchronize()
        pass_ms[form] = e0.elapsed_time(e1) / 3
        runners[form] = (r, cfg_f)
        log(f"{form} form: stage-1 shape {cfg_f}, {pass_ms[form]:.3f} ms a pass")
    form = min(pass_ms, key=pass_ms.get)
    br, cfg = runners[form]
    fused = form == "fused"
    log(f"form: {form}")

    def step(ev=None):
        br.replay(ev)
        return br.collect(), (None if fused else
                              br.C[:L1 * br.plan.ldcs[0]].view(L1, br.plan.ldcs[0])[:, :L2])

    for _ in range(args.warmup):
        res, Cb = step()
    torch.cuda.synchronize()
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        res, Cb = step(events[k])
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / args.steps
    cost_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in events]))
    s1_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in events]))
    s2_ms = float(np.mean([e[2].elapsed_time(e[3]) for e in events]))
    cells = plan.cells

    # FP64 plain form on the same pair: the reference for cost error and agreement.
    if Cb is None:
        # the fused form's cost values are those of the BF16 cost kernel
        # (bit-identical, tested); build that matrix here for the error figure
        Cb = cost_matrix_from_features_gpu(F1d, F2d, dtype=torch.bfloat16)
    C64 = cost_matrix_from_features_gpu(F1d, F2d)
    cerr = 0.0
    for r0 in range(0, L1, 4096):
        cerr = max(cerr, float((Cb[r0:r0 + 4096].double() - C64[r0:r0 + 4096]).abs().max()))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bc64, wp64 = PG.parflex(C64, STEPS, WEIGHTS, BETA, L, profile_dir=None)
    torch.cuda.synchronize()
    plain_ms = (time.perf_counter() - t0) * 1e3
    bc = float(res.best_cost[0])
    agree = path_agreement(wp64, res.paths[0], tol=2)
    log(f"precision: {ms_step:.2f} ms/step (cost {cost_ms:.2f}, stage1 {s1_ms:.2f}, stage2 "
        f"{s2_ms:.2f}); plain FP64 {plain_ms:.1f} ms; agreement {agree:.5f}; max |dC| {cerr:.3g}")
    out = _base(args, round(cells / (ms_step * 1e-3) / 1e9, 3), "GCUPS", ms_step,
                "bf16 cost, f32 DP, f64 stage 2",
                {"workload": args.workload, "L1": L1, "L2": L2, "K": int(F1.shape[0]), "L": L,
                 "tiles": plan.ntiles, "cells": cells, "steps": STEPS, "weights": WEIGHTS,
                 "beta": BETA,
                 "form": "precision: BF16 MFMA cost " + ("computed inside stage 1 (fused)"
                                                       if fused else "materialised")
                         + ", FP32 DP, FP64 stage 2"})
    out["data"] = "synthetic (seeded ReLU-Gaussian pair with known time warp, features in HBM)"
    out["dp_gcups"] = round(cells / (s1_ms * 1e-3) / 1e9, 3)
    out["pairs_per_s"] = round(1e3 / ms_step, 3)
    out["form_pass_ms"] = {k: round(v, 4) for k, v in pass_ms.items()}
    out["cost_ms"] = round(cost_ms, 4)   # fused: the BF16 feature conversion only
    if not fused:
        out["cost_tflops"] = round(2.0 * F1.shape[0] * L1 * L2 / (cost_ms * 1e-3) / 1e12, 2)
    out["stage1_ms"] = round(s1_ms, 4)
    out["stage2_ms"] = round(s2_ms, 4)
    out["plain_gpu"] = {"ms": round(plain_ms, 3), "pairs_per_s": round(1e3 / plain_ms, 3)}
    out["accuracy"] = {"cost_max_abs_err": cerr, "best_cost": bc, "best_cost_fp64": bc64,
                       "best_cost_abs_err": abs(bc - bc64),
                       "best_cost_rel_err": abs(bc - bc64) / abs(bc64),
                       "frames_within_2": round(agree, 6), "reference": "plain FP64 path"}
    out["stage1_shape"] = list(cfg)
    if fused:
        out["roofline"] = _roofline(cells, s1_ms, 0.25, _feat_kernel(cfg, L, K=128))
    else:
        out["roofline"] = _roofline(cells, s1_ms, 2.25,
                                    "k_stage1<2,__hip_bfloat16,float,false,true,false,1024>")
    if not args.no_cpu:
        C = C64.cpu().numpy()
        # single thread: the top-left 16000 x 16000 block (2.6e8 cells, 16 tiles)
        C1 = np.ascontiguousarray(C[:16000, :16000])
        out["cpu_baseline"], exact = _cpu_baseline(
            args, [C], [C1], L, f"the whole {L1}x{L2} pair", lambda cps: cps / 1e9, "GCUPS",
            expect=[(bc64, wp64)])
        out["cpu_baseline"]["single_thread"]["sample"] += " (top-left 16000x16000 block)"
        out["parity"] = {"plain_gpu_vs_cpu_oracle_exact": bool(exact)}
    return out


def main():
    args = _parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "path (stage 1 sharded by tile bands) is not implemented.", file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.set_device(0)
    if args.copy_probe:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy probe: 1073741824 bytes read, 1073741824 bytes written")
        del src, dst
    if args.workload.startswith("chroma_learned"):
        out = bench_learned(args, torch)
    elif args.workload.startswith("chroma_batch"):
        out = bench_batch(args, torch)
    elif args.workload.startswith("relu"):
        out = bench_precision(args, torch)
    else:
        out = bench_plain(args, torch)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
