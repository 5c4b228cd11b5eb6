"""Write golden input/output vectors from the reference implementation.

Run by hand in the development container only (the reference checkout is not
available anywhere else):

    python tests/golden/make_golden.py [--reference /root/reference] [--big]

The reference's modules import numba at module level and numba is not
installed, so a stand-in module whose ``jit``/``njit`` decorators return the
function unchanged is registered first; the reference then runs as plain
Python/NumPy (IEEE double, no FMA contraction), which is the arithmetic its
Numba kernels compile to with fastmath off.

Only data is written (``*.npz`` next to this script): seeded inputs and the
reference's outputs for them.  Cases:
  flexdtw_seed1     FlexDTW.flexdtw on the seeds/shapes of the reference test
                    tests/test_gpu_flexdtw.py:20-39
  tiles_seed2       run_flexdtw_on_tiles + stage 2 (tests/test_gpu_flexdtw.py:42-62)
  edges_seed4       _build_edge_data + _build_valid_mask (tests/test_gpu_flexdtw.py:65-89)
  parflex_seed3     parflex on tests/test_gpu_flexdtw.py:92-106 inputs
  stage2_*          multi-tile stage 2 on synthetic chroma pairs and random C
                    (ragged edge tiles, several L, weights and step sets)
  cfg0 (--big)      the 2000 x 2400, L=500 chroma workload
  stage2_chroma400_L5_cap
                    a 400 x 440 chroma pair at L=5 (100 x 110 tiles): the
                    stitched path crosses more than 101 tiles, so the
                    reference's 101-tile stitch cap (Parflex.py:1041-1049)
                    truncates it
  stage2_sync_random
                    (--sync) sync_tile_overlap_edges (Parflex.py:919-969) on
                    random D_arr / L_arr with infinite slots and near-equal
                    corners, where every pass changes values (on real stage-2
                    arrays the fill already leaves the overlaps consistent)
  stage2_presync    (--presync) propagate_tile_edge_costs outputs before the
                    overlap sync (Parflex.py:862-912), D_arr and L_arr, for
                    every stage-2 case above
"""
from __future__ import annotations

import argparse
import os
import sys
import time
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))


def _install_numba_standin():
    mod = types.ModuleType("numba")

    def _deco(*args, **kwargs):
        if len(args) == 1 and callable(args[0]) and not kwargs:
            return args[0]
        return lambda f: f

    mod.jit = _deco
    mod.njit = _deco
    sys.modules["numba"] = mod


def _stage2_fields(r):
    out = {
        "s2_D_arr": r["D_arr"], "s2_L_arr": r["L_arr"], "s2_edge_lens": r["edge_lens"],
        "s2_best_cost": np.float64(r["best_cost"]),
        "s2_best_end": np.array(r["best_end"], dtype=np.int64),
        "s2_stitched_wp": np.asarray(r["stitched_wp"], dtype=np.int64).reshape(-1, 2),
        "s2_top_D": r["edge_summary"]["top"]["D"], "s2_top_L": r["edge_summary"]["top"]["L"],
        "s2_right_D": r["edge_summary"]["right"]["D"],
        "s2_right_L": r["edge_summary"]["right"]["L"],
    }
    segs = []
    for (name, seg_idx), meta in sorted(r["best_per_segment"].items()):
        segs.append([0 if name == "top" else 1, seg_idx, meta["chunk_i"], meta["chunk_j"],
                     meta["edge"], meta["idx"], meta["global_coord"][0],
                     meta["global_coord"][1]])
    out["s2_segments"] = np.array(segs, dtype=np.int64).reshape(-1, 8)
    out["s2_segment_norm"] = np.array(
        [meta["norm_cost"] for _, meta in sorted(r["best_per_segment"].items())],
        dtype=np.float64)
    return out


def _tiles_fields(chunks, n1, n2):
    out = {}
    for (i, j), ch in chunks.items():
        out[f"tile_{i}_{j}_D"] = ch["D"]
        out[f"tile_{i}_{j}_S"] = ch["S"]
        out[f"tile_{i}_{j}_B"] = ch["B"]
        out[f"tile_{i}_{j}_best"] = np.float64(ch["best_cost"])
        out[f"tile_{i}_{j}_wp"] = np.asarray(ch["wp"], dtype=np.int64)
    out["n1"] = np.int64(n1)
    out["n2"] = np.int64(n2)
    return out


def _stage2_case(P, C, L, steps, weights, beta):
    chunks, L_out, n1, n2 = P.run_flexdtw_on_tiles(C, L=L, steps=steps, weights=weights,
                                                   buffer=1, profile_dir=None)
    tiled = P.build_tiled_metadata(chunks, L_out, n1, n2, C,
                                   stage1_params={"steps": np.array(steps).reshape(-1, 2),
                                                  "weights": np.array(weights),
                                                  "buffer": 1.0})
    tiled["chunks_dict"] = chunks
    out = {"C": C, "L": np.int64(L), "steps": np.array(steps, dtype=np.int64).reshape(-1, 2),
           "weights": np.array(weights, dtype=np.float64), "beta": np.float64(beta)}
    try:
        r = P.run_stage2_from_tiled(tiled, C, beta=beta, profile_dir=None,
                                    backtrace_segments=True)
        out.update(_stage2_fields(r))
        out["found"] = np.int64(1)
        segp = []
        for key, val in sorted(r["paths_per_segment"].items()):
            pts = np.asarray(val["path"], dtype=np.int64).reshape(-1, 2)
            segp.append(pts)
        out["s2_segment_path_lens"] = np.array([len(p) for p in segp], dtype=np.int64)
        out["s2_segment_paths"] = (np.concatenate(segp) if segp else
                                   np.zeros((0, 2), np.int64))
        bc, wp = P.parflex(C, steps, weights, beta, L=L, profile_dir=None)
        out["parflex_best_cost"] = np.float64(bc)
        out["parflex_wp"] = np.asarray(wp, dtype=np.int64)
    except ValueError as exc:
        if "No valid endpoint" not in str(exc):
            raise
        out["found"] = np.int64(0)
    return out, chunks, n1, n2


def _stage2_cases(steps, weights):
    """Multi-tile stage 2 on synthetic chroma pairs and random matrices.
    Chroma cases store the features; C is rebuilt with cost_matrix_det,
    whose summation order is fixed, so the fixture stays small."""
    from parflexdtw_amd.synth import make_pair
    cases = []
    pairs = {300: make_pair(300, 360, seed=101), 420: make_pair(420, 380, seed=102),
             250: make_pair(250, 300, seed=103), 600: make_pair(600, 700, seed=104)}
    cases.append(("stage2_chroma300_L50", pairs[300], 50, steps, weights, 0.1))
    cases.append(("stage2_chroma300_L64", pairs[300], 64, steps, [1.25, 3, 3], 0.1))
    cases.append(("stage2_chroma420_L37", pairs[420], 37, steps, weights, 0.1))
    cases.append(("stage2_chroma250_L400", pairs[250], 400, steps, weights, 0.1))
    rng = np.random.RandomState(7)
    cases.append(("stage2_rand_L9", rng.rand(61, 53), 9, steps, weights, 0.1))
    cases.append(("stage2_rand_L5", rng.rand(20, 23), 5, steps, weights, 0.3))
    cases.append(("stage2_rand_4step", rng.rand(90, 110), 25,
                  [(1, 1), (1, 2), (2, 1), (2, 2)], [2, 3, 3, 4], 0.1))
    cases.append(("stage2_nocand", rng.rand(30, 30), 8, steps, weights, 1.0))
    cases.append(("stage2_chroma600_L100", pairs[600], 100, steps, weights, 0.1))
    # 100 x 110 tiles of 5 x 5: the path crosses more than 101 tiles
    cases.append(("stage2_chroma400_L5_cap", make_pair(400, 440, seed=105), 5, steps, weights,
                  0.1))
    return cases


def _case_C(src):
    from parflexdtw_amd.synth import cost_matrix_det
    if isinstance(src, tuple):
        F1, F2, _ = src
        return np.ascontiguousarray(cost_matrix_det(F1, F2)), F1, F2
    return np.ascontiguousarray(src), None, None


def _write_cases(P, cases):
    for name, src, L, st, wt, beta in cases:
        t0 = time.time()
        C, F1, F2 = _case_C(src)
        o, _, _, _ = _stage2_case(P, C, L, st, wt, beta)
        if F1 is not None:
            del o["C"]
            o["F1"] = F1
            o["F2"] = F2
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **o)
        print(f"{name} written ({time.time() - t0:.1f}s)", flush=True)


def _write_presync(P, only):
    """D_arr / L_arr of propagate_tile_edge_costs (before the overlap sync)
    for every stage-2 fixture, from the inputs stored in the fixture."""
    names = ["tiles_seed2", "parflex_seed3"] + [c[0] for c in _stage2_cases([(1, 1)], [1])]
    if only is not None:
        names = [n for n in names if n in only]
    out = {}
    for name in names:
        with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
            g = {k: z[k] for k in z.files}
        if "C" in g:
            C = g["C"]
        else:
            from parflexdtw_amd.synth import cost_matrix_det
            C = cost_matrix_det(g["F1"], g["F2"])
        L = int(g["L"])
        chunks, L_out, n1, n2 = P.run_flexdtw_on_tiles(C, L=L, steps=g["steps"],
                                                       weights=g["weights"], buffer=1,
                                                       profile_dir=None)
        D_arr, L_arr, edge_lens = P.propagate_tile_edge_costs(chunks, L_out, n1, n2,
                                                              buffer_param=1, profile_dir=None)
        out[f"{name}__D_arr"] = D_arr.copy()
        out[f"{name}__L_arr"] = L_arr.copy()
        out[f"{name}__edge_lens"] = np.asarray(edge_lens, dtype=np.int64)
        # the sync of the same arrays must give the stored stage-2 result
        D2, L2s = P.sync_tile_overlap_edges(D_arr.copy(), L_arr.copy(), edge_lens, n1, n2)
        if "s2_D_arr" in g:
            assert np.array_equal(D2, g["s2_D_arr"]) and np.array_equal(L2s, g["s2_L_arr"])
        print(f"presync {name} done", flush=True)
    np.savez_compressed(os.path.join(HERE, "stage2_presync.npz"), **out)
    print("stage2_presync written", flush=True)


def _write_sync_random(P):
    """Random inputs for the overlap sync: finite values, +inf slots with
    probability 0.4, and corner pairs that differ by 0, 5e-11 or 2e-10 (the
    1e-10 threshold of Parflex.py:965 on both sides)."""
    out = {}
    specs = [(50, 61, 9, 0), (23, 40, 6, 1), (100, 80, 17, 2), (9, 9, 5, 3)]
    for k, (L1, L2, L, seed) in enumerate(specs):
        rng = np.random.RandomState(500 + seed)
        hop = L - 1
        n1, n2 = -(-(L1 - 1) // hop), -(-(L2 - 1) // hop)
        D = np.full((n1, n2, 2, L), np.inf)
        La = np.full((n1, n2, 2, L), np.inf)
        el = np.zeros((n1, n2, 2), np.int64)
        for i in range(n1):
            for j in range(n2):
                rows = min(i * hop + L, L1) - i * hop
                cols = min(j * hop + L, L2) - j * hop
                el[i, j] = (cols, rows)
                for e, n in ((0, cols), (1, rows)):
                    v = rng.rand(n) * 10.0
                    ln = rng.randint(1, 200, size=n).astype(np.float64)
                    dead = rng.rand(n) < 0.4
                    v[dead] = np.inf
                    ln[dead] = np.inf
                    D[i, j, e, :n] = v
                    La[i, j, e, :n] = ln
                if rng.rand() < 0.5 and np.isfinite(D[i, j, 0, cols - 1]):
                    delta = [0.0, 5e-11, -5e-11, 2e-10, -2e-10][rng.randint(5)]
                    D[i, j, 1, rows - 1] = D[i, j, 0, cols - 1] + delta
        out[f"c{k}_D_in"] = D.copy()
        out[f"c{k}_L_in"] = La.copy()
        out[f"c{k}_edge_lens"] = el
        out[f"c{k}_shape"] = np.array([L1, L2, L, n1, n2], np.int64)
        Do, Lo = P.sync_tile_overlap_edges(D, La, el, n1, n2)
        out[f"c{k}_D_out"] = np.asarray(Do).copy()
        out[f"c{k}_L_out"] = np.asarray(Lo).copy()
        changed = int(np.sum(~((out[f"c{k}_D_in"] == out[f"c{k}_D_out"]) |
                               (np.isinf(out[f"c{k}_D_in"]) & np.isinf(out[f"c{k}_D_out"])))))
        print(f"sync case {k}: {changed} slots changed", flush=True)
    out["ncases"] = np.int64(len(specs))
    np.savez_compressed(os.path.join(HERE, "stage2_sync_random.npz"), **out)
    print("stage2_sync_random written", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--big", action="store_true", help="also write the 2000x2400 workload")
    ap.add_argument("--only", default=None,
                    help="comma-separated fixture names: write only these")
    ap.add_argument("--sync", action="store_true",
                    help="write stage2_sync_random.npz (sync on random arrays)")
    ap.add_argument("--presync", action="store_true",
                    help="write stage2_presync.npz from the inputs of the stage-2 fixtures")
    args = ap.parse_args()
    only = None if args.only is None else set(args.only.split(","))

    _install_numba_standin()
    sys.path.insert(0, args.reference)
    import FlexDTW  # noqa: E402
    import Parflex as P  # noqa: E402

    from parflexdtw_amd.synth import cost_matrix_det, make_pair  # noqa: E402

    steps = [(1, 1), (1, 2), (2, 1)]
    weights = [2, 3, 3]

    if args.presync:
        _write_presync(P, only)
        return
    if args.sync:
        _write_sync_random(P)
        return
    if only is not None:
        _write_cases(P, [c for c in _stage2_cases(steps, weights) if c[0] in only])
        return

    # FlexDTW.flexdtw single tiles (tests/test_gpu_flexdtw.py:20-39).
    rng = np.random.RandomState(1)
    out = {}
    for k, (H, W) in enumerate([(5, 6), (17, 19), (48, 50)]):
        C = np.abs(rng.randn(H, W).astype(np.float64)) * 0.1
        bc, wp, D, Pm, B, debug = FlexDTW.flexdtw(C, steps=steps, weights=weights, buffer=1)
        out[f"C{k}"] = C
        out[f"best{k}"] = np.float64(bc)
        out[f"wp{k}"] = np.asarray(wp, dtype=np.int64)
        out[f"D{k}"] = D
        out[f"P{k}"] = Pm
        out[f"B{k}"] = B
        out[f"debug{k}"] = np.array([[d[0], d[1], d[2], d[3]] for d in debug], dtype=np.float64)
    # Extra single tiles: other weights, a 4-step set, a larger random tile.
    extra = [((1, 1), (1, 2), (2, 1)), ((1, 1), (1, 2), (2, 1), (1, 0)),
             ((1, 1), (2, 2), (1, 2), (2, 1))]
    ew = [[1.25, 3.0, 3.0], [2.0, 3.0, 3.0, 4.0], [2.0, 4.0, 3.0, 3.0]]
    rng = np.random.RandomState(11)
    for k, (st, wt) in enumerate(zip(extra, ew)):
        C = rng.rand(37, 41)
        bc, wp, D, Pm, B, debug = FlexDTW.flexdtw(C, steps=np.array(st), weights=wt, buffer=1)
        out[f"xC{k}"] = C
        out[f"xsteps{k}"] = np.array(st, dtype=np.int64)
        out[f"xweights{k}"] = np.array(wt, dtype=np.float64)
        out[f"xbest{k}"] = np.float64(bc)
        out[f"xwp{k}"] = np.asarray(wp, dtype=np.int64)
        out[f"xD{k}"] = D
        out[f"xP{k}"] = Pm
        out[f"xB{k}"] = B
    np.savez_compressed(os.path.join(HERE, "flexdtw_seed1.npz"), **out)
    print("flexdtw_seed1 written", flush=True)

    # Tiled stage 1 + stage 2 (tests/test_gpu_flexdtw.py:42-62).
    rng = np.random.RandomState(2)
    C = np.abs(rng.randn(40, 44).astype(np.float64)) * 0.08
    o, chunks, n1, n2 = _stage2_case(P, C, 14, steps, weights, 0.1)
    o.update(_tiles_fields(chunks, n1, n2))
    np.savez_compressed(os.path.join(HERE, "tiles_seed2.npz"), **o)
    print("tiles_seed2 written", flush=True)

    # Edge tensors and valid mask (tests/test_gpu_flexdtw.py:65-89).
    rng = np.random.RandomState(4)
    C = np.abs(rng.randn(24, 28).astype(np.float64)) * 0.06
    L = 11
    chunks, _, n1, n2 = P.run_flexdtw_on_tiles(C, L=L, steps=steps, weights=weights,
                                               profile_dir=None)
    t = P._build_edge_data(chunks, n1, n2, L)
    vm = P._build_valid_mask(chunks, n1, n2, L)
    names = ["edge_Df", "edge_start_r", "edge_start_c", "edge_Cf_olap", "edge_lens",
             "chunk_rows", "chunk_cols", "chunk_bounds", "chunk_valid"]
    o = {n: a for n, a in zip(names, t)}
    o["valid_mask"] = vm
    o["C"] = C
    o["L"] = np.int64(L)
    np.savez_compressed(os.path.join(HERE, "edges_seed4.npz"), **o)
    print("edges_seed4 written", flush=True)

    # parflex (tests/test_gpu_flexdtw.py:92-106 inputs).
    rng = np.random.RandomState(3)
    C = np.abs(rng.randn(16, 18).astype(np.float64)) * 0.07
    o, _, _, _ = _stage2_case(P, C, 10, steps, weights, 0.1)
    np.savez_compressed(os.path.join(HERE, "parflex_seed3.npz"), **o)
    print("parflex_seed3 written", flush=True)

    _write_cases(P, _stage2_cases(steps, weights))

    if args.big:
        F1, F2, _ = make_pair(2000, 2400, seed=0)
        C = cost_matrix_det(F1, F2)
        t0 = time.time()
        bc, wp = P.parflex(C, steps, weights, 0.1, L=500, profile_dir=None)
        np.savez_compressed(os.path.join(HERE, "cfg0_chroma2000_L500.npz"),
                            F1=F1, F2=F2, L=np.int64(500), beta=np.float64(0.1),
                            steps=np.array(steps, dtype=np.int64),
                            weights=np.array(weights, dtype=np.float64),
                            parflex_best_cost=np.float64(bc),
                            parflex_wp=np.asarray(wp, dtype=np.int64))
        print(f"cfg0 written ({time.time() - t0:.1f}s)", flush=True)


if __name__ == "__main__":
    main()
