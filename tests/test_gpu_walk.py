"""The stage-2 back-pointer walk (k_s2_walk, under backtrace_and_stitch,
Parflex.py:976-1103) on back-pointers that no DP would write: cells set to an
invalid code and cells set to random valid codes, so that walks stop on an
invalid code or on a step off the tile in the middle of their fast steps
(above the start row), far from the start cell.  The device stitch from many
edge slots must equal the oracle's backtrace_and_stitch over the same codes
point for point."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STD = [(1, 1), (1, 2), (2, 1)]
CASES = [  # (C shape, L, steps, weights, fraction invalid, fraction random valid, seed)
    ((600, 700), 256, STD, [2, 3, 3], 0.002, 0.0, 1),
    ((600, 700), 256, STD, [2, 3, 3], 0.0, 0.05, 2),
    ((700, 520), 300, STD, [2, 3, 3], 0.001, 0.02, 3),
    # two steps: codes 2 and 3 are invalid
    ((500, 640), 200, [(1, 1), (1, 2)], [2, 3], 0.003, 0.03, 4),
    # four steps: every code is a step; (1, 0) walks straight up to row 0
    ((520, 600), 256, [(1, 0), (1, 1), (1, 2), (2, 1)], [1, 2, 3, 3], 0.0, 0.08, 5),
    # tiles of more than 64 columns per block and 2 x 2 tiles of 1024
    ((1500, 1300), 1024, STD, [2, 3, 3], 0.0005, 0.01, 6),
]


@pytest.fixture(scope="module")
def torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.mark.parametrize("case", range(len(CASES)))
def test_walk_on_corrupted_back_pointers_vs_oracle(torch_gpu, oracle_mod, case):
    torch = torch_gpu
    from parflexdtw_amd import engine
    from parflexdtw_amd._plan import BatchPlan
    shape, L, steps, weights, f_inv, f_rand, seed = CASES[case]
    K = len(steps)
    rng = np.random.RandomState(700 + seed)
    C = rng.rand(*shape)
    plan = BatchPlan([C.shape], L)
    sw = engine.StepWeights(steps, weights)
    C_dev = torch.from_numpy(C).cuda()
    ws = plan.allocate(C_dev.device)
    engine.stage1(C_dev, plan, sw, ws)
    s1 = oracle_mod.run_stage1(C, L, steps, weights)
    n1, n2 = plan.grids[0]
    Bp = ws["Bp"].cpu().numpy().view(np.uint32).copy()
    B = s1["B"].copy()
    for i in range(n1):
        for j in range(n2):
            t = plan.tiles[i * n2 + j]
            R, W, boff = int(t[2]), int(t[3]), int(t[4])
            nw = (R + 15) // 16
            words = Bp[boff:boff + nw * W].reshape(nw, W)
            r = np.arange(R)
            codes = ((words[r // 16, :] >> (2 * (r % 16))[:, None].astype(np.uint32)) & 3)
            codes = codes.astype(np.int64)
            ob = B[i, j].reshape(-1)[:R * W].reshape(R, W)
            assert np.array_equal(codes, ob.astype(np.int64)), (i, j)
            u = rng.rand(R, W)
            codes = np.where(u < f_inv, 3, codes)
            codes = np.where((u >= f_inv) & (u < f_inv + f_rand), rng.randint(0, K, size=(R, W)),
                             codes)
            # repack: 16 rows per word, row r at bits 2 (r % 16)
            pad = np.zeros((nw * 16, W), np.uint64)
            pad[:R] = codes
            packed = np.zeros((nw, W), np.uint64)
            for k in range(16):
                packed |= pad[k::16] << np.uint64(2 * k)
            Bp[boff:boff + nw * W] = packed.astype(np.uint32).reshape(-1)
            B[i, j].reshape(-1)[:R * W] = codes.reshape(-1).astype(np.int8)
    ws["Bp"].copy_(torch.from_numpy(Bp.view(np.int32)))   # the packed words are int32
    s1c = dict(s1)
    s1c["B"] = B
    # end slots: every edge of every tile at a few positions (first, last, random)
    reqs = []
    for i in range(n1):
        for j in range(n2):
            t = plan.tiles[i * n2 + j]
            R, W = int(t[2]), int(t[3])
            for et, n in ((0, W), (1, R)):
                for idx in sorted({0, n - 1, int(rng.randint(0, n)), int(rng.randint(0, n))}):
                    reqs.append((0, i, j, et, idx))
    got = engine.stitch_requests(plan, sw, ws, reqs)
    torch.cuda.synchronize()
    for q, (_, i, j, et, idx) in enumerate(reqs):
        want = oracle_mod.backtrace_and_stitch(s1c, i, j, et, idx)
        # within the path capacity, so that no truncation rule is involved
        assert len(want) < shape[0] + shape[1] + 8
        assert np.array_equal(got[q], want), (case, i, j, et, idx, got[q].shape, want.shape)
