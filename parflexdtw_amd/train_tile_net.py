"""Train and calibrate the tile network on the GPU.

    python -m parflexdtw_amd.train_tile_net --out parflexdtw_amd/data/tile_net_L2000.npz

Training pairs are seeded synthetic chroma pairs (synth.batch_pairs, seeds
40000+p); their labels are the paths of the plain FP64 method (the batched
FP64 form, bit-identical to parflex) at the tile length L: a pooled
32 x 32 block is positive when the path crosses it.  The network is trained
with the project's kernels only (tile_net.TileNet: MFMA forward and backward,
weighted BCE, Adam with FP32 master weights).

Calibration on held-out pairs (seeds 60000+p) fixes
  * threshold: the lowest score of any tile within `guard` frames of a
    validation path, minus a safety margin (so no validation pair falls back
    for its selection); a tile's score is its largest logit over its pooled
    cells extended by `ext` cells;
  * cost_bound: 1.25 x the largest normalised cost of a plain path seen in
    training and validation (a learned result above it falls back).
Everything is saved with the weights.
"""
from __future__ import annotations

import argparse
import math
import sys
import time

import numpy as np

STEPS = [(1, 1), (1, 2), (2, 1)]
WEIGHTS = [2.0, 3.0, 3.0]


def plain_paths(pairs, L, beta=0.1, cell_budget=3.0e9, log=None):
    """Best costs and paths of the plain FP64 method, in sub-batches that keep
    the FP64 cost matrices within cell_budget cells."""
    import torch
    from . import engine
    out_cost, out_path = [], []
    i = 0
    while i < len(pairs):
        j, cells = i, 0
        while j < len(pairs) and (j == i or cells + pairs[j][0].shape[1] * pairs[j][1].shape[1]
                                  <= cell_budget):
            cells += pairs[j][0].shape[1] * pairs[j][1].shape[1]
            j += 1
        br = engine.BatchRunner(pairs[i:j], L, STEPS, WEIGHTS, beta, torch.float64)
        br.launch()
        r = br.collect()
        out_cost.extend(r.best_cost.tolist())
        out_path.extend(r.paths)
        del br
        torch.cuda.empty_cache()
        if log:
            log(f"plain paths {j}/{len(pairs)}")
        i = j
    return np.array(out_cost), out_path


def build_dataset(net, pairs, paths):
    """Per pair: pooled features on the device (the network input is rebuilt
    from them at every step), the pooled shape and the block labels."""
    import torch
    from .tile_net import path_block_labels
    data = []
    for (F1, F2), wp in zip(pairs, paths):
        A, B, H, W, K = net.pooled_features(torch.from_numpy(F1).cuda(),
                                            torch.from_numpy(F2).cuda())
        lab = path_block_labels(wp, F1.shape[1], F2.shape[1])
        data.append((A, B, H, W, K, torch.from_numpy(lab.reshape(-1)).cuda(), int(lab.sum())))
    return data


def train(net, data, epochs=20, lr=2e-3, batch=8, pos_weight=20.0, seed=0, log=None,
          lr_end=None):
    """Adam over shuffled images, `batch` images per step, per-image mean
    loss, learning rate decaying from lr to lr_end (default 0) on a cosine;
    returns the mean loss of every epoch."""
    import torch
    rng = np.random.default_rng(seed)
    hist = []
    nsteps = epochs * -(-len(data) // batch)
    step = 0
    for ep in range(epochs):
        order = rng.permutation(len(data))
        net.loss.zero_()
        for b0 in range(0, len(order), batch):
            idx = order[b0:b0 + batch]
            for k in idx:
                A, B, H, W, K, lab, _ = data[k]
                X, _ = net.input_from_pooled(A, B, H, W, K)
                acts, Z = net.forward(X, H, W)
                net.backward(acts, Z, H, W, lab, pos_weight, 1.0 / (H * W * len(idx)))
            # cosine decay
            end = 0.0 if lr_end is None else lr_end
            cur = end + (lr - end) * 0.5 * (1.0 + math.cos(math.pi * step / max(nsteps, 1)))
            net.adam(lr=cur)
            step += 1
        torch.cuda.synchronize()
        hist.append(float(net.loss.item()) * batch / len(data))
        if log:
            log(f"epoch {ep + 1}/{epochs}: loss {hist[-1]:.4f}")
    return hist


def calibrate(net, pairs, paths, L, ext=8, guard=128, safety=1.0):
    """Threshold = min over the validation pairs of the score of every tile
    within `guard` frames of the path, minus `safety` (so no validation pair
    falls back for its selection); also the fraction of tiles selected."""
    import torch
    from .learned import path_tiles
    from .tile_net import select_tiles
    mins, fr = [], []
    scores = []
    for (F1, F2), wp in zip(pairs, paths):
        s, n1, n2 = net.score_pair(torch.from_numpy(F1).cuda(), torch.from_numpy(F2).cuda(), L,
                                   ext)
        s = s.cpu().numpy().reshape(n1, n2)
        scores.append(s)
        need = path_tiles(wp, L, n1, n2, guard, F1.shape[1], F2.shape[1])
        mins.append(float(s.reshape(-1)[need].min()))
    thr = float(min(mins)) - safety
    for s in scores:
        fr.append(select_tiles(s, s.shape[0], s.shape[1], thr).mean())
    return thr, float(np.mean(fr)), np.array(mins)


def _gen_pair(args):
    """Pair seed_base+p of synth.batch_pairs (run in a worker process)."""
    seed, n_lo, n_hi = args
    from .synth import make_pair
    rng = np.random.default_rng(seed)
    N = int(rng.integers(n_lo, n_hi + 1))
    M = int(round(N * rng.uniform(0.8, 1.25)))
    F1, F2, _ = make_pair(N, M, seed=seed)
    return F1, F2


def labelled_set(net, n, seed_base, n_lo, n_hi, L, chunk=256, log=None, keep_pairs=False,
                 workers=16):
    """Generate n pairs in chunks (host worker processes), label them with the
    plain FP64 method and keep only their pooled features and block labels
    on the device (the host features of a chunk are dropped after use)."""
    import multiprocessing
    import os
    from concurrent.futures import ProcessPoolExecutor
    data, costs, kept, kept_paths = [], [], [], []
    nw = max(1, min(workers, os.cpu_count() or 1))
    with ProcessPoolExecutor(nw, mp_context=multiprocessing.get_context("spawn")) as pool:
        for c0 in range(0, n, chunk):
            todo = [(seed_base + p, n_lo, n_hi) for p in range(c0, min(n, c0 + chunk))]
            pairs = list(pool.map(_gen_pair, todo, chunksize=4))
            data_c, costs_c, pairs_c, paths_c = _label_chunk(net, pairs, L)
            data.extend(data_c)
            costs.extend(costs_c)
            if keep_pairs:
                kept.extend(pairs_c)
                kept_paths.extend(paths_c)
            if log:
                log(f"labelled {min(n, c0 + chunk)}/{n} pairs (seeds {seed_base}+)")
    return data, np.array(costs), kept, kept_paths


def _label_chunk(net, pairs, L):
    cost, paths = plain_paths(pairs, L)
    return build_dataset(net, pairs, paths), cost.tolist(), pairs, paths


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--pairs", type=int, default=10000)
    ap.add_argument("--val", type=int, default=64)
    ap.add_argument("--n-lo", type=int, default=2000)
    ap.add_argument("--n-hi", type=int, default=30000)
    ap.add_argument("--L", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--pos-weight", type=float, default=20.0)
    ap.add_argument("--ext", type=int, default=8, help="selection margin, pooled cells")
    ap.add_argument("--guard", type=int, default=128, help="fallback guard, frames")
    args = ap.parse_args(argv)

    from .tile_net import TileNet

    def log(m):
        print(f"[train] {m}", file=sys.stderr, flush=True)

    net = TileNet(seed=0)
    t0 = time.perf_counter()
    data, tr_cost, _, _ = labelled_set(net, args.pairs, 40000, args.n_lo, args.n_hi, args.L,
                                       log=log)
    _, va_cost, va, va_paths = labelled_set(net, args.val, 60000, args.n_lo, args.n_hi, args.L,
                                            log=log, keep_pairs=True)
    log(f"data and plain labels in {time.perf_counter() - t0:.1f} s")
    pos = sum(d[6] for d in data)
    npx = sum(d[2] * d[3] for d in data)
    log(f"dataset: {npx} pooled cells, {pos} on a path ({pos / npx:.4f})")
    cost_bound = 1.25 * float(np.nanmax(np.concatenate([tr_cost, va_cost])))
    meta = dict(ext=args.ext, guard=args.guard, cost_bound=cost_bound, L=args.L, pool=32,
                train_pairs=args.pairs, n_lo=args.n_lo, n_hi=args.n_hi, epochs=args.epochs,
                lr=args.lr, pos_weight=args.pos_weight)
    t0 = time.perf_counter()
    hist = []
    for ep in range(args.epochs):
        # one epoch at a time so that every epoch leaves a usable checkpoint
        def sched(e):
            return args.lr * 0.5 * (1 + math.cos(math.pi * e / args.epochs))
        h = train(net, data, 1, sched(ep), pos_weight=args.pos_weight, seed=ep,
                  lr_end=sched(ep + 1))
        hist.extend(h)
        thr, frac, mins = calibrate(net, va, va_paths, args.L, args.ext, args.guard)
        log(f"epoch {ep + 1}/{args.epochs}: loss {h[-1]:.5f}, threshold {thr:.3f}, "
            f"validation tiles selected {frac:.4f} ({time.perf_counter() - t0:.0f} s)")
        net.save(args.out, threshold=thr, loss_history=np.array(hist),
                 val_selected_fraction=frac, val_min_logits=mins,
                 train_seconds=time.perf_counter() - t0, epochs_done=ep + 1, **meta)
    log(f"saved {args.out}; cost bound {cost_bound:.5f}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
