"""GPU tests of the learned pipeline: what the trained network gains over the
method without it, checked against the plain FP64 method, and that training with
the project's kernels lowers the loss."""
import os

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

STEPS = [(1, 1), (1, 2), (2, 1)]
WEIGHTS = [2, 3, 3]
MODEL = os.path.join(ROOT, "parflexdtw_amd", "data", "tile_net_L2000.npz")


@pytest.fixture(scope="module")
def torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_trained_network_gains_over_full_tiles(torch_gpu):
    """Held-out pairs (seeds outside training and validation): the trained
    network's selection computes well under half of the cells, needs no
    fallback and keeps >= 99.5% of frames within +-2 of the plain FP64
    path; an untrained network selects nothing useful, so every pair falls
    back to the full method (no gain, same answer)."""
    from parflexdtw_amd.learned import LearnedRunner
    from parflexdtw_amd.synth import batch_pairs, path_agreement
    from parflexdtw_amd.tile_net import TileNet
    from parflexdtw_amd.train_tile_net import plain_paths
    net, meta = TileNet.load(MODEL)
    L = int(meta["L"])
    pairs = [(a, b) for a, b, _ in batch_pairs(6, 95000, 14000, 24000)]
    cost, paths = plain_paths(pairs, L)
    args = (float(meta["threshold"]), int(meta["ext"]), int(meta["guard"]),
            float(meta["cost_bound"]))
    res, st = LearnedRunner(pairs, L, STEPS, WEIGHTS, 0.1, net, *args).run()
    frac = st["cells_computed"] / st["cells_full"]
    assert frac < 0.5, frac
    assert st["fallback_pairs"] == []
    for p in range(len(pairs)):
        assert path_agreement(paths[p], res.paths[p], tol=2) >= 0.995, p
        assert abs(res.best_cost[p] - cost[p]) <= 2e-2 * cost[p]
    res_u, st_u = LearnedRunner(pairs, L, STEPS, WEIGHTS, 0.1, TileNet(seed=1), *args).run()
    assert st_u["fallback_pairs"] == list(range(len(pairs)))
    for p in range(len(pairs)):
        assert path_agreement(paths[p], res_u.paths[p], tol=2) >= 0.995, p


def test_training_lowers_the_loss(torch_gpu):
    from parflexdtw_amd.synth import batch_pairs
    from parflexdtw_amd.tile_net import TileNet
    from parflexdtw_amd.train_tile_net import build_dataset, plain_paths, train
    pairs = [(a, b) for a, b, _ in batch_pairs(16, 70000, 2000, 5000)]
    _, paths = plain_paths(pairs, 500)
    net = TileNet(seed=4)
    data = build_dataset(net, pairs, paths)
    hist = train(net, data, epochs=12, lr=3e-3, batch=4, pos_weight=20.0)
    assert hist[-1] < 0.6 * hist[0], hist


def test_scoring_graph_replays_equal_eager(torch_gpu):
    """The scoring of all pairs replayed from its HIP graph equals the eager
    pass, replay after replay, with eager work in between."""
    torch = torch_gpu
    from parflexdtw_amd.learned import LearnedRunner
    from parflexdtw_amd.synth import batch_pairs
    from parflexdtw_amd.tile_net import TileNet
    net, meta = TileNet.load(MODEL)
    pairs = [(a, b) for a, b, _ in batch_pairs(4, 96000, 5000, 9000)]
    eager = LearnedRunner(pairs, 2000, STEPS, WEIGHTS, 0.1, net, 0.0, graph=False).score()
    lg = LearnedRunner(pairs, 2000, STEPS, WEIGHTS, 0.1, net, 0.0, graph=True)
    for r in range(3):
        got = lg.score()
        torch.randn(1 << 20, device="cuda").sum().item()   # unrelated eager work
        lg._scores.fill_(-1.0)
        for a, b in zip(eager, got):
            assert np.array_equal(a, b), r


def test_batched_scoring_equals_per_pair_scoring(torch_gpu):
    """The runner's scoring (batched pooling and statistics, the first layer
    from pooled features, pairs over side streams) gives each pair the tile
    logits of TileNet.score_pair on that pair alone, bit for bit."""
    torch = torch_gpu
    from parflexdtw_amd.learned import LearnedRunner
    from parflexdtw_amd.synth import batch_pairs
    from parflexdtw_amd.tile_net import TileNet
    net, meta = TileNet.load(MODEL)
    pairs = [(a, b) for a, b, _ in batch_pairs(7, 98000, 800, 7000)]
    L, ext = 600, 5
    got = LearnedRunner(pairs, L, STEPS, WEIGHTS, 0.1, net, 0.0, ext=ext, guard=100,
                        graph=False).score()
    for p, (F1, F2) in enumerate(pairs):
        s, n1, n2 = net.score_pair(torch.from_numpy(F1).cuda(), torch.from_numpy(F2).cuda(), L,
                                   ext)
        assert np.array_equal(got[p], s.cpu().numpy().reshape(n1, n2)), p


def test_fused_network_equals_per_layer_kernels(torch_gpu):
    """pf_tile_net (the four layers in one launch over every pair) gives the
    logit maps of the per-layer kernels bit for bit, for maps from 2 x 3
    pooled cells up to several job rows (128) and job columns (58), and
    pf_tile_scores_batch the per-pair tile scores."""
    from parflexdtw_amd.learned import LearnedRunner
    from parflexdtw_amd.synth import make_pair
    from parflexdtw_amd.tile_net import TileNet
    net, meta = TileNet.load(MODEL)
    sizes = [(60, 75), (100, 40), (1900, 1850), (4130, 4400), (9000, 7777), (33, 33),
             (5000, 1200)]
    pairs = [make_pair(n, m, seed=4242 + q)[:2] for q, (n, m) in enumerate(sizes)]
    kw = dict(ext=3, guard=10, graph=False)
    fused = LearnedRunner(pairs, 300, STEPS, WEIGHTS, 0.1, net, 0.0, fused_net=True, **kw)
    layer = LearnedRunner(pairs, 300, STEPS, WEIGHTS, 0.1, net, 0.0, fused_net=False, **kw)
    a, b = fused.logit_maps(), layer.logit_maps()
    for p in range(len(pairs)):
        assert a[p].shape == b[p].shape
        assert np.array_equal(a[p], b[p]), (p, float(np.abs(a[p] - b[p]).max()))
    for p, (x, y) in enumerate(zip(fused.score(), layer.score())):
        assert np.array_equal(x, y), p


def test_device_selection_equals_host_selection(torch_gpu):
    """pf_select_tiles (the selection, tile tables and diagonal lists on the
    device) gives the host selection's tables exactly: the same selection,
    full and work tables, counts, and per tile diagonal the same selected
    and pruned tiles (their order inside a diagonal is free), at several
    thresholds including none and every tile selected, and with the
    scores' exact threshold ties."""
    torch = torch_gpu
    from parflexdtw_amd import engine
    from parflexdtw_amd.learned import LearnedRunner
    from parflexdtw_amd.synth import make_pair
    from parflexdtw_amd.tile_net import TileNet
    net, _ = TileNet.load(MODEL)
    sizes = [(3000, 2800), (700, 900), (4100, 2500), (60, 75), (2600, 4000)]
    pairs = [make_pair(n, m, seed=5151 + q)[:2] for q, (n, m) in enumerate(sizes)]
    lr = LearnedRunner(pairs, 300, STEPS, WEIGHTS, 0.1, net, 0.0, ext=3, guard=10, graph=False)
    plan = lr.plan
    d_scores = lr._score_dev()
    flat = d_scores.cpu().numpy()
    ties = float(np.sort(flat)[len(flat) // 3])
    for thr in (-np.inf, float(np.median(flat)), ties, 1e9, float(flat.min())):
        lr.threshold = thr
        nsel, b_words, cells, h_diag = lr._select_device(d_scores)
        sel, full, work, _, _, bw_h, _, cells_h = lr._pruned_tables(flat >= thr)
        t = lr._sel
        assert nsel == len(work) == int(sel.sum()), thr
        assert b_words == bw_h and cells == cells_h
        assert np.array_equal(t["sel"].cpu().numpy().astype(bool), sel)
        assert np.array_equal(t["full"].cpu().numpy(), full)
        assert np.array_equal(t["work"].cpu().numpy()[:nsel], work)
        lst_h, hd_h = engine.diagonal_lists(plan, sel)
        assert np.array_equal(h_diag, hd_h)
        lst = t["list"].cpu().numpy()
        for d in range(plan.max_diag):
            o, nl, npr = (int(v) for v in hd_h[d])
            assert sorted(lst[o:o + nl]) == sorted(lst_h[o:o + nl])
            assert sorted(lst[o + nl:o + nl + npr]) == sorted(lst_h[o + nl:o + nl + npr])


def test_selection_check_kernel_matches_host_rule(torch_gpu):
    """pf_selection_check flags a pair exactly when a tile within `guard`
    frames of its stitched path is unselected (host rule: learned.path_tiles)."""
    torch = torch_gpu
    from parflexdtw_amd._lib import check, lib
    from parflexdtw_amd.engine import BatchRunner
    from parflexdtw_amd.learned import path_tiles
    from parflexdtw_amd.synth import batch_pairs
    pairs = [(a, b) for a, b, _ in batch_pairs(5, 97000, 1500, 4000)]
    L = 300
    br = BatchRunner(pairs, L, STEPS, WEIGHTS, 0.1, torch.bfloat16)
    br.launch()
    res = br.collect()
    rng = np.random.default_rng(0)
    for trial in range(6):
        guard = [0, 5, 40, 100, 149, 7][trial]
        sel = rng.random(br.plan.ntiles) < [0.97, 0.9, 0.99, 0.8, 0.995, 1.0][trial]
        flag = torch.zeros(len(pairs), dtype=torch.int32, device="cuda")
        d_sel = torch.from_numpy(sel.astype(np.uint8)).cuda()
        mp = max(F1.shape[1] + F2.shape[1] for F1, F2 in pairs) + 8
        check(lib().pf_selection_check(br.ws["pairs"].data_ptr(), len(pairs), mp,
                                       br.ws["path"].data_ptr(), br.ws["path_len"].data_ptr(),
                                       d_sel.data_ptr(), L, guard, flag.data_ptr(),
                                       torch.cuda.current_stream().cuda_stream), "check")
        got = flag.cpu().numpy()
        for p, (F1, F2) in enumerate(pairs):
            n1, n2 = br.plan.grids[p]
            base = int(br.plan.pairs[p, 4])
            s = sel[base:base + n1 * n2]
            want = not s[path_tiles(res.paths[p], L, n1, n2, guard, F1.shape[1],
                                    F2.shape[1])].all()
            assert bool(got[p]) == want, (trial, p)
