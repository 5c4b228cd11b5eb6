"""The BASELINE workloads at their full sizes on the GPU.

* 16000 x ~19000 is in test_gpu_parity.py;
* 256 chroma pairs, N ~ U[3000, 20000]: the tuned fused BF16 batched form
  against the plain FP64 form (sub-batches; a sample of small pairs checked
  against the oracle);
* one relu K = 128 pair of 40000 x ~36500 frames at L = 4000: the plain FP64
  form equals the oracle, the precision form (BF16 cost, FP32 DP, FP64 stage
  2) agrees with it and its cost error is bounded;
* the 128 held-out chroma pairs of N ~ U[20000, 60000] at L = 2000 through
  the learned form, against the plain FP64 form, and one pair above 2^31
  cells index-exact against the oracle (64-bit indexing; the reference's
  flat index overflows there, gpu_flexdtw.py:188).

Each test prints what it runs before it runs it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = [(1, 1), (1, 2), (2, 1)]
WEIGHTS = [2, 3, 3]


@pytest.fixture(scope="module")
def torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _say(msg):
    print(msg, flush=True)


def test_batch256_tuned_fused_vs_plain_fp64(torch_gpu, oracle_mod):
    torch = torch_gpu
    from parflexdtw_amd import engine
    from parflexdtw_amd.synth import batch_pairs, cost_matrix_det, path_agreement
    from parflexdtw_amd.train_tile_net import plain_paths
    trip = batch_pairs(256, 2000, 3000, 20000)
    pairs = [(a, b) for a, b, _ in trip]
    cells = sum(a.shape[1] * b.shape[1] for a, b in pairs)
    _say(f"batch256: {cells / 1e9:.1f} G matrix cells; tuning L in 500..4000 (fused BF16)")
    br, report = engine.tune_batch(pairs, [500, 1000, 2000, 4000], STEPS, WEIGHTS, 0.1,
                                   torch.bfloat16, fused=True)
    assert sorted(report) == [500, 1000, 2000, 4000]
    br.launch()
    res = br.collect()
    L = br.plan.L
    _say(f"batch256: tuned L={L}, config={br.ws.get('s1_config')}; plain FP64 at L={L}")
    cost, paths = plain_paths(pairs, L, cell_budget=4.0e9)
    agree = np.array([path_agreement(paths[p], res.paths[p], tol=2) for p in range(len(pairs))])
    assert res.found.all()
    assert agree.min() >= 0.995, (int(agree.argmin()), float(agree.min()))
    rel = np.abs(res.best_cost - cost) / cost
    assert rel.max() < 2e-2, float(rel.max())
    # the plain FP64 batched paths of the smallest pairs equal the oracle's
    order = np.argsort([a.shape[1] * b.shape[1] for a, b in pairs])[:4]
    for p in order:
        F1, F2 = pairs[p]
        bo, wpo = oracle_mod.parflex(cost_matrix_det(F1, F2), STEPS, WEIGHTS, 0.1, L,
                                     nthreads=16)
        assert bo == cost[p]
        assert np.array_equal(wpo, paths[p]), p


def test_precision_relu128_40000_vs_oracle(torch_gpu, oracle_mod):
    torch = torch_gpu
    from parflexdtw_amd import Parflex_gpu, gpu_flexdtw
    from parflexdtw_amd import engine
    from parflexdtw_amd.synth import cfg_pair, path_agreement
    F1, F2, _ = cfg_pair(3)
    N, M = F1.shape[1], F2.shape[1]
    L = 4000
    _say(f"precision: relu K=128, {N} x {M}, L={L}: FP64 cost and path")
    C64 = gpu_flexdtw.cost_matrix_from_features_gpu(F1, F2)
    bc, wp = Parflex_gpu.parflex(C64, STEPS, WEIGHTS, 0.1, L, profile_dir=None)
    # BF16 cost error against the FP64 cost (on the device)
    C16 = gpu_flexdtw.cost_matrix_from_features_gpu(F1, F2, dtype=torch.bfloat16)
    err = 0.0
    for r0 in range(0, N, 4096):
        d = (C16[r0:r0 + 4096].double() - C64[r0:r0 + 4096]).abs().max().item()
        err = max(err, d)
    del C16
    _say(f"precision: max |C_bf16 - C_f64| = {err:.3e}; fused form")
    assert err < 1.6e-2
    br = engine.BatchRunner([(F1, F2)], L, STEPS, WEIGHTS, 0.1, torch.bfloat16, fused=True)
    br.autotune()
    br.launch()
    r = br.collect()
    assert r.found[0]
    agree = path_agreement(wp, r.paths[0], tol=2)
    _say(f"precision: agreement {agree:.5f}, best cost {r.best_cost[0]:.8f} vs {bc:.8f}")
    assert agree >= 0.995
    assert abs(r.best_cost[0] - bc) <= 1e-2 * bc
    _say("precision: oracle on the FP64 cost (16 threads)")
    C = C64.cpu().numpy()
    del C64
    torch.cuda.empty_cache()
    bo, wpo = oracle_mod.parflex(C, STEPS, WEIGHTS, 0.1, L, nthreads=16)
    assert bc == bo
    assert np.array_equal(wp, wpo)


def test_learned_heldout128_vs_plain_fp64(torch_gpu):
    import os
    from conftest import ROOT
    from parflexdtw_amd.learned import LearnedRunner
    from parflexdtw_amd.synth import batch_pairs, path_agreement
    from parflexdtw_amd.tile_net import TileNet
    from parflexdtw_amd.train_tile_net import plain_paths
    net, meta = TileNet.load(os.path.join(ROOT, "parflexdtw_amd", "data", "tile_net_L2000.npz"))
    L = int(meta["L"])
    _say("learned: 128 held-out pairs (seeds 90000+), N ~ U[20000, 60000], L=2000")
    trip = batch_pairs(128, 90000, 20000, 60000)
    pairs = [(a, b) for a, b, _ in trip]
    lr = LearnedRunner(pairs, L, STEPS, WEIGHTS, 0.1, net, float(meta["threshold"]),
                       int(meta["ext"]), int(meta["guard"]), float(meta["cost_bound"]))
    res, st = lr.run()
    _say(f"learned: {st['selected_tiles']} of {st['tiles']} tiles, "
         f"fallbacks {st['fallback_pairs']}; plain FP64 paths")
    assert st["cells_computed"] < 0.5 * st["cells_full"]
    cost, paths = plain_paths(pairs, L, cell_budget=8.0e9)
    agree = np.array([path_agreement(paths[p], res.paths[p], tol=2) for p in range(len(pairs))])
    n_frames = np.array([paths[p].shape[1] for p in range(len(pairs))])
    total = float((agree * n_frames).sum() / n_frames.sum())
    _say(f"learned: frame agreement {total:.5f} (min per pair {agree.min():.5f})")
    assert res.found.all()
    assert total >= 0.995
    assert agree.min() >= 0.99
    assert (np.abs(res.best_cost - cost) / cost).max() < 2e-2


def test_pair_above_2_31_cells_vs_oracle(torch_gpu, oracle_mod):
    """48000 x 50862 chroma pair (2.44e9 cells) at L = 2000: element offsets
    of C, back-pointer words and path buffers pass 2^31; the plain FP64 form
    equals the oracle index for index."""
    torch = torch_gpu
    from parflexdtw_amd import Parflex_gpu, gpu_flexdtw
    from parflexdtw_amd.synth import make_pair
    F1, F2, _ = make_pair(48000, 56000, seed=90777)
    N, M = F1.shape[1], F2.shape[1]
    assert N * M > 2 ** 31
    _say(f"2^31: {N} x {M} = {N * M / 1e9:.2f} G cells, L=2000, FP64 on the GPU")
    C_dev = gpu_flexdtw.cost_matrix_from_features_gpu(F1, F2)
    bc, wp = Parflex_gpu.parflex(C_dev, STEPS, WEIGHTS, 0.1, 2000, profile_dir=None)
    _say("2^31: copy C to the host, oracle (16 threads)")
    C = C_dev.cpu().numpy()
    del C_dev
    torch.cuda.empty_cache()
    bo, wpo = oracle_mod.parflex(C, STEPS, WEIGHTS, 0.1, 2000, nthreads=16)
    assert bc == bo
    assert np.array_equal(wp, wpo)
