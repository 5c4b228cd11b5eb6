"""GPU tests of the batched alignment evaluator (pf_eval_paths).

Its errors must equal the host evaluator's, np.interp(x, rows*hop, cols*hop)
- y as eval_tools.py:44-46 and 05_Evaluate_Benchmarks_L.ipynb#cell8 compute
them, bit for bit, on every branch of np.interp (before the first point,
after the last, on a sample point, between points, NaN and infinite times,
one-point paths, repeated query frames), and its error rates must equal the
per-threshold means of #cell23 exactly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HOP = 512 / 22050


@pytest.fixture(scope="module")
def torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _walk(n, rng, steps, start=(0, 0)):
    st = np.asarray(steps, dtype=np.int64)
    d = st[rng.integers(0, len(st), size=max(n - 1, 0))]
    pts = np.concatenate([np.zeros((1, 2), np.int64), np.cumsum(d, axis=0)]) + np.asarray(start)
    return np.ascontiguousarray(pts.T)


def _times(wp, rng, hop, k):
    """k query times over the path's span and beyond, every path point's own
    time, and NaN / +-inf; reference times random."""
    if wp is None or wp.shape[1] == 0:
        x = rng.uniform(0, 100, k)
    else:
        lo, hi = wp[0, 0] * hop, wp[0, -1] * hop
        span = max(hi - lo, hop)
        x = rng.uniform(lo - 0.05 * span - 3 * hop, hi + 0.05 * span + 3 * hop, k)
        x = np.concatenate([x, wp[0, rng.integers(0, wp.shape[1], 40)] * hop,
                            [lo, hi, np.nextafter(hi, -np.inf), np.nextafter(lo, np.inf)]])
    x = np.concatenate([x, [np.nan, np.inf, -np.inf]])
    rng.shuffle(x)
    y = rng.uniform(0, 100, len(x))
    return x, y


def _host_errors(wp, x, y, hop):
    from parflexdtw_amd.evaluate import _path_errors
    if wp is None or wp.shape[1] == 0:
        return np.full(len(x), np.inf) - y
    return _path_errors(wp, x, y, hop)


@pytest.mark.parametrize("hop,maxTol", [(HOP, 1000), (1.0, 4095), (HOP, 0)])
def test_eval_paths_equal_host_evaluator(torch_gpu, hop, maxTol):
    from parflexdtw_amd.evaluate import error_rates, evaluate_paths
    rng = np.random.default_rng(11)
    S = [(1, 1), (1, 2), (2, 1)]
    paths = [_walk(1, rng, S), _walk(2, rng, S), _walk(5, rng, S), _walk(300, rng, S),
             _walk(6000, rng, S, start=(17, 40)), None, np.zeros((2, 0), np.int64),
             _walk(700, rng, [(0, 1), (1, 0), (1, 1)], start=(3, 0)),
             _walk(129, rng, S, start=(1000, 5))]
    gts = [_times(wp, rng, hop, k) for wp, k in
           zip(paths, [20, 30, 50, 600, 5000, 10, 7, 900, 256])]
    print(f"evaluate_paths hop={hop} maxTol={maxTol} queries={len(paths)}", flush=True)
    got = evaluate_paths(paths, gts, hop, maxTol)
    want = [_host_errors(wp, x, y, hop) for wp, (x, y) in zip(paths, gts)]
    for q in range(len(paths)):
        assert got.errs[q].shape == want[q].shape, q
        assert np.array_equal(got.errs[q], want[q], equal_nan=True), q
    rates, flat = error_rates(want, maxTol)
    assert got.beats == len(flat) == int(got.hist.sum())
    assert np.array_equal(got.rates, rates)


def test_eval_batch_reads_device_paths(torch_gpu, tmp_path):
    """evaluate_batch on a BatchRunner's device paths equals the evaluator on
    the collected host paths; align_batch_to_files writes those paths."""
    from parflexdtw_amd.engine import BatchRunner
    from parflexdtw_amd.evaluate import (align_batch_to_files, evaluate_batch, evaluate_paths,
                                         loadAlignment)
    from parflexdtw_amd.synth import beat_times, make_pair
    torch = torch_gpu
    shapes = [(700, 800), (1100, 950), (433, 521), (1500, 1720), (260, 301)]
    trip = [make_pair(n, m, seed=70 + i) for i, (n, m) in enumerate(shapes)]
    pairs = [(a, b) for a, b, _ in trip]
    gts = [beat_times(w, HOP, every=7, offset=i) for i, (_, _, w) in enumerate(trip)]
    steps, weights = [(1, 1), (1, 2), (2, 1)], [1.25, 3.0, 3.0]
    print("evaluate_batch on 5 ragged pairs, L=256", flush=True)
    br = BatchRunner(pairs, 256, steps, weights, beta=0.1, dtype=torch.float64)
    br.launch()
    got = evaluate_batch(br, gts, HOP, 1000)
    res = br.collect()
    ref = evaluate_paths(res.paths, gts, HOP, 1000)
    for q in range(len(pairs)):
        assert np.array_equal(got.errs[q], ref.errs[q]), q
        assert np.array_equal(got.errs[q], _host_errors(res.paths[q], *gts[q], HOP)), q
    assert np.array_equal(got.hist, ref.hist) and np.array_equal(got.rates, ref.rates)
    # the alignments are good: most beats within 100 ms
    assert got.rates[100] < 0.05
    files = [tmp_path / f"q{i}.pkl" for i in range(len(pairs))]
    res2 = align_batch_to_files(pairs, files, L=256, steps=steps, weights=weights)
    for q in range(len(pairs)):
        assert np.array_equal(loadAlignment(files[q]), res.paths[q]), q
        assert res2.best_cost[q] == res.best_cost[q]


def test_eval_known_answers_on_device(torch_gpu, tmp_path):
    """unit_test_eval.ipynb#cell62-63 through the device evaluator."""
    import test_evaluate_host as H
    from parflexdtw_amd.evaluate import evaluate_paths, getTimestamps
    wp = H._unit_wp().astype(np.int64)
    paths, gts, names = [], [], []
    for b in H.BENCHMARKS:
        for case in ("exact_match", "shifted_2"):
            a1, a2 = H._annotations(b, case)
            H._write_beat(str(tmp_path / b / case / "a1.beat"), a1)
            H._write_beat(str(tmp_path / b / case / "a2.beat"), a2)
            gts.append(getTimestamps(tmp_path / b / case / "a1.beat",
                                     tmp_path / b / case / "a2.beat"))
            paths.append(wp)
            names.append((b, case))
    got = evaluate_paths(paths, gts, HOP, 1000)
    expect = {"matching": 0, "subseq10": 0, "post5": 0, "partialOverlap": 41, "pre5": 5,
              "prepost5": 5}
    for (b, case), e in zip(names, got.errs):
        a = np.abs(e)
        if case == "exact_match":
            assert np.all(a == expect[b]), (b, case)
        elif expect[b] == 0:
            assert np.all(a == 2), (b, case)
        else:
            assert np.all(a == expect[b] - 2) or np.all(a == expect[b] + 2), (b, case)


def test_align_system_gpu_writes_drop_in_result(torch_gpu, tmp_path):
    from parflexdtw_amd import Parflex_gpu as PG
    from parflexdtw_amd.evaluate import align_system_gpu, get_outfile, loadAlignment
    from parflexdtw_amd.gpu_flexdtw import cost_matrix_from_features_gpu
    from parflexdtw_amd.synth import make_pair
    F1, F2, _ = make_pair(900, 1000, seed=5)
    f = get_outfile(tmp_path, "matching", "gpu_parflex", "a__b")
    bc, wp = align_system_gpu(F1, F2, f, L=300)
    bc2, wp2 = PG.parflex(cost_matrix_from_features_gpu(F1, F2),
                          np.array([1, 1, 1, 2, 2, 1]).reshape(-1, 2),
                          np.array([1.25, 3.0, 3.0]), 0.1, 300, profile_dir=None)
    assert bc == bc2 and np.array_equal(wp, wp2)
    assert np.array_equal(loadAlignment(f), wp)
