"""GPU tests of the batched pipeline (features -> cost -> stage 1 -> stage 2
for many ragged pairs, HIP-graph replay, stage-1 autotuner) and of the
precision form (BF16 MFMA cost, FP32 DP, FP64 stage 2).

The FP64 batched form is checked bit for bit against the oracle pair by pair.
The precision form has no exact counterpart: its best cost is compared with
the FP64 path's (relative error bound in the test) and its path with the FP64
path frame by frame (at least 99.5% of F1 frames within +-2 F2 frames)."""
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


def _pairs(shapes, seed0, K=12, kind="chroma"):
    from parflexdtw_amd.synth import make_pair
    out = []
    for p, (n, m) in enumerate(shapes):
        F1, F2, _ = make_pair(n, m, K=K, seed=seed0 + p, kind=kind)
        out.append((F1, F2))
    return out


RAGGED = [(700, 800), (1100, 950), (433, 521), (1500, 1720), (260, 301)]


def test_batch_runner_f64_matches_oracle(torch_gpu, oracle_mod):
    from parflexdtw_amd.engine import BatchRunner
    from parflexdtw_amd.synth import cost_matrix_det
    torch = torch_gpu
    pairs = _pairs(RAGGED, 500)
    br = BatchRunner(pairs, 256, STEPS, WEIGHTS, beta=0.1, dtype=torch.float64)
    br.launch()
    res = br.collect()
    for p, (F1, F2) in enumerate(pairs):
        bc, wp = oracle_mod.parflex(cost_matrix_det(F1, F2), STEPS, WEIGHTS, 0.1, 256)
        assert res.found[p]
        assert res.best_cost[p] == bc, p
        assert np.array_equal(res.paths[p], wp), p


@pytest.mark.parametrize("dtype", ["float64", "bfloat16"])
def test_batch_graph_replay_equals_eager(torch_gpu, dtype):
    """Three replays of the captured cost + stage 1 + stage 2 graph, with the
    outputs overwritten by eager work before each, all equal the eager run."""
    from parflexdtw_amd.engine import BatchRunner
    torch = torch_gpu
    pairs = _pairs(RAGGED, 600)
    br = BatchRunner(pairs, 300, STEPS, WEIGHTS, beta=0.1, dtype=getattr(torch, dtype))
    br.launch()
    ref = br.collect()
    br.capture()
    for r in range(3):
        br.ws["path"].fill_(-5)
        br.ws["best_f"].fill_(7.0)
        br.ws["Bp"].fill_(0)
        br.C.fill_(0)
        torch.cuda.synchronize()
        br.replay()
        torch.cuda.synchronize()
        got = br.collect()
        assert np.array_equal(got.found, ref.found), r
        assert np.array_equal(got.best_cost, ref.best_cost, equal_nan=True), r
        for p in range(len(pairs)):
            assert np.array_equal(got.paths[p], ref.paths[p]), (r, p)


def test_autotune_keeps_results(torch_gpu):
    """Every stage-1 launch shape the autotuner tries gives the same FP64 result."""
    from parflexdtw_amd import engine
    torch = torch_gpu
    pairs = _pairs(RAGGED[:3], 700)
    br = engine.BatchRunner(pairs, 600, STEPS, WEIGHTS, beta=0.1, dtype=torch.float64)
    br.launch()
    ref = br.collect()
    engine._TUNED.clear()
    best, times = br.autotune()
    assert best in times and len(times) >= 3
    for cfg in times:
        engine.stage1(br.C, br.plan, br.sw, br.ws, config=cfg)
        engine.stage2(br.C, br.plan, br.sw, br.ws)
        got = br.collect()
        assert np.array_equal(got.best_cost, ref.best_cost), cfg
        for p in range(len(pairs)):
            assert np.array_equal(got.paths[p], ref.paths[p]), (cfg, p)


@pytest.mark.parametrize("kind,K,N,M,L", [("relu", 128, 3000, 3300, 1000),
                                          ("chroma", 12, 2000, 2400, 500),
                                          ("relu", 128, 1200, 1000, 257)])
def test_precision_form_agreement(torch_gpu, kind, K, N, M, L):
    """BF16 MFMA cost + FP32 DP + FP64 stage 2 against the FP64 plain path."""
    from parflexdtw_amd import Parflex_gpu as PG
    from parflexdtw_amd import gpu_flexdtw as G
    from parflexdtw_amd.synth import make_pair, path_agreement
    torch = torch_gpu
    F1, F2, _ = make_pair(N, M, K=K, seed=N + K, kind=kind)
    C64 = G.cost_matrix_from_features_gpu(F1, F2)
    bc64, wp64 = PG.parflex(C64, STEPS, WEIGHTS, 0.1, L, profile_dir=None)
    for dt in (torch.bfloat16, torch.float32):
        Cp = G.cost_matrix_from_features_gpu(F1, F2, dtype=dt)
        bcp, wpp = PG.parflex(Cp, STEPS, WEIGHTS, 0.1, L, profile_dir=None)
        assert path_agreement(wp64, wpp, tol=2) >= 0.995, dt
        assert abs(bcp - bc64) <= 2e-2 * abs(bc64) + 1e-3, (dt, bcp, bc64)


@pytest.mark.parametrize("dtype", ["float64", "float32", "bfloat16"])
def test_batched_cost_equals_single(torch_gpu, dtype):
    """One pf_cost_batch launch over ragged pairs writes, for every pair,
    exactly what pf_cost writes for that pair alone; FP64 also equals the
    fixed-order host cost."""
    from parflexdtw_amd.engine import BatchRunner
    from parflexdtw_amd.gpu_flexdtw import cost_matrix_from_features_gpu
    from parflexdtw_amd.synth import cost_matrix_det
    torch = torch_gpu
    dt = getattr(torch, dtype)
    pairs = _pairs([(300, 257), (129, 1000), (64, 64), (2, 3), (700, 129)], 800, K=20,
                   kind="relu")
    br = BatchRunner(pairs, 50, STEPS, WEIGHTS, beta=0.1, dtype=dt)
    br.cost()
    torch.cuda.synchronize()
    for p, (F1, F2) in enumerate(pairs):
        N, M = F1.shape[1], F2.shape[1]
        off, ldc = int(br.plan.pairs[p, 5]), int(br.plan.pairs[p, 6])
        assert ldc % 8 == 0 and off % 8 == 0 and ldc >= M
        got = br.C[off:off + N * ldc].view(N, ldc)[:, :M]
        one = cost_matrix_from_features_gpu(F1, F2, dtype=dt)
        assert torch.equal(got, one), p
        if dtype == "float64":
            assert np.array_equal(got.cpu().numpy(), cost_matrix_det(F1, F2)), p


@pytest.mark.parametrize("K,kind,L,config", [(12, "chroma", 300, None), (12, "chroma", 700, (3, 2, 16)),
                                             (20, "relu", 257, (2, 1, 32)), (12, "chroma", 1000, (1, 8, 32)),
                                             (128, "relu", 520, (2, 2, 7)),
                                             # one strip of 4 columns per lane (k_sweep4): 1, 4,
                                             # 8 and 16 waves (idle waves past narrow tiles)
                                             (12, "chroma", 232, (1, 1, 8)),
                                             (12, "chroma", 1000, (1, 4, 32)),
                                             (12, "chroma", 2000, (1, 8, 64)),
                                             (12, "chroma", 4000, (1, 16, 64)),
                                             (16, "relu", 700, (1, 4, 16)),
                                             # K = 20 and 128: chains of 2 and 8
                                             # 16x16x16 MFMAs per block
                                             (20, "relu", 250, (1, 1, 8)),
                                             (128, "relu", 520, (1, 4, 16)),
                                             (128, "relu", 2000, (1, 8, 64))])
def test_fused_stage1_equals_materialised_bf16(torch_gpu, K, kind, L, config):
    """Stage 1 computing C inside the sweep (pf_stage1_features) gives the
    same back-pointers, edges and paths, bit for bit, as stage 1 reading the
    materialised BF16 cost matrix (same MFMA operands and rounding)."""
    from parflexdtw_amd import engine
    from parflexdtw_amd.engine import BatchRunner
    torch = torch_gpu
    pairs = _pairs([(700, 800), (433, 521), (1100, 950), (260, 301)], 900 + K, K=K, kind=kind)
    ref = BatchRunner(pairs, L, STEPS, WEIGHTS, beta=0.1, dtype=torch.bfloat16)
    fus = BatchRunner(pairs, L, STEPS, WEIGHTS, beta=0.1, dtype=torch.bfloat16, fused=True)
    for br in (ref, fus):
        # the materialised form has no 16-wave strips of 4 columns per lane:
        # it runs its default shape there (the results do not depend on it)
        if config is not None and (br is fus or config[1] * 64 * 4 <= 2048):
            br.ws["s1_config"] = config
        br.launch()
    torch.cuda.synchronize()
    assert torch.equal(ref.ws["Bp"], fus.ws["Bp"])
    # edge slots that exist (position < edge length); the rest are never written
    tiles = ref.plan.tiles
    valid = np.zeros((len(tiles), 2, L), bool)
    for t, row in enumerate(tiles):
        valid[t, 0, :row[3]] = True
        valid[t, 1, :row[2]] = True
    valid = valid.reshape(-1)
    for k in ("edge_D", "edge_P", "edge_C"):
        assert np.array_equal(ref.ws[k].cpu().numpy()[valid], fus.ws[k].cpu().numpy()[valid]), k
    a, b = ref.collect(), fus.collect()
    assert np.array_equal(a.best_cost, b.best_cost, equal_nan=True)
    for p in range(len(pairs)):
        assert np.array_equal(a.paths[p], b.paths[p]), p
    assert fus.C.numel() == 1   # no cost matrix is stored


ONEHOT_CONFIGS = [  # (N, M, L, launch shape)
    (150, 2100, 2048, (1, 8, 64)), (140, 1100, 1000, (1, 4, 16)), (150, 600, 232, (1, 1, 8)),
    (3, 700, 700, (1, 4, 16)), (18, 300, 300, (1, 2, 16)), (131, 3000, 2100, (1, 16, 64)),
]


@pytest.mark.parametrize("case", range(len(ONEHOT_CONFIGS)))
def test_fused_stage1_one_hot_features_vs_oracle(torch_gpu, oracle_mod, case):
    """The features form on one-hot frames: C = 1 - a.b is exactly 0 or 1, and
    with tiles of at most 150 rows every FP32 sum and cross product of the DP
    is exact, so the edges and every packed back-pointer equal the FP64
    oracle's on that C (one strip of 4 columns per lane with 1 to 16 waves;
    tiles of 3 and 18 rows; ragged last tiles)."""
    from parflexdtw_amd.engine import BatchRunner
    torch = torch_gpu
    N, M, L, cfg = ONEHOT_CONFIGS[case]
    rng = np.random.default_rng(500 + case)
    a, b = rng.integers(0, 12, N), rng.integers(0, 12, M)
    F1 = np.zeros((12, N))
    F2 = np.zeros((12, M))
    F1[a, np.arange(N)] = 1.0
    F2[b, np.arange(M)] = 1.0
    C = (a[:, None] != b[None, :]).astype(np.float64)
    br = BatchRunner([(F1, F2)], L, STEPS, WEIGHTS, beta=0.1, dtype=torch.bfloat16, fused=True)
    br.ws["s1_config"] = cfg
    br.launch()
    torch.cuda.synchronize()
    s1 = oracle_mod.run_stage1(C, L, STEPS, WEIGHTS)
    n1, n2 = br.plan.grids[0]
    eD = br.ws["edge_D"].cpu().numpy().reshape(n1, n2, 2, L)
    eP = br.ws["edge_P"].cpu().numpy().reshape(n1, n2, 2, L)
    eC = br.ws["edge_C"].cpu().numpy().reshape(n1, n2, 2, L)
    Bp = br.ws["Bp"].cpu().numpy().view(np.uint32)
    for i in range(n1):
        for j in range(n2):
            t = br.plan.tiles[i * n2 + j]
            R, W, boff = int(t[2]), int(t[3]), int(t[4])
            r0, c0 = i * (L - 1), j * (L - 1)
            assert np.array_equal(eC[i, j, 0, :W], C[r0, c0:c0 + W]), (i, j)
            assert np.array_equal(eC[i, j, 1, :R], C[r0:r0 + R, c0]), (i, j)
            assert np.array_equal(eD[i, j, 0, :W], s1["edge_D"][i, j, 0, :W]), (i, j)
            assert np.array_equal(eD[i, j, 1, :R], s1["edge_D"][i, j, 1, :R]), (i, j)
            assert np.array_equal(eP[i, j, 0, :W], s1["edge_P"][i, j, 0, :W]), (i, j)
            assert np.array_equal(eP[i, j, 1, :R], s1["edge_P"][i, j, 1, :R]), (i, j)
            words = Bp[boff:boff + ((R + 15) // 16) * W].reshape(-1, W)
            r = np.arange(R)
            codes = (words[r // 16, :] >> (2 * (r % 16))[:, None].astype(np.uint32)) & 3
            ob = s1["B"][i, j].reshape(-1)[:R * W].reshape(R, W)
            assert np.array_equal(codes.astype(np.int8), ob), (i, j)


def test_stage2_reads_no_edges_of_tiles_marked_without_stage1(torch_gpu):
    """A tile table entry with column 0 < 0 (a tile whose stage 1 did not run)
    makes stage 2 treat the tile's edges as D = +inf, P = 0, C = 0 without
    reading them: the same outputs as filling those edges with +inf, 0, 0,
    whatever the unread slots hold.  D of a marked tile is written only in
    the slots some step reads (position 0 and the corners; the last row when
    the tile above is unmarked or the tile is in the top row; the last
    column when the tile to the right is unmarked or the tile is in the right
    column); the other slots keep a finite poison value, which a read would
    carry into the paths."""
    from parflexdtw_amd import engine
    from parflexdtw_amd.engine import BatchRunner
    torch = torch_gpu
    pairs = _pairs([(1500, 1700), (900, 1250), (2100, 1900)], 4242)
    L = 300
    br = BatchRunner(pairs, L, STEPS, WEIGHTS, beta=0.1, dtype=torch.bfloat16)
    br.launch()
    torch.cuda.synchronize()
    plan, ws = br.plan, br.ws
    rng = np.random.default_rng(5)
    pruned = rng.random(plan.ntiles) < 0.4
    slots = np.repeat(pruned, 2 * L)
    idx = torch.from_numpy(np.nonzero(slots)[0]).cuda()
    outs = []
    for marked in (False, True):
        w = dict(ws)
        for k in ("edge_D", "edge_P", "edge_C"):
            w[k] = ws[k].clone()
        if marked:
            # garbage in the slots that must not be read
            w["edge_D"][idx] = -7.0
            w["edge_P"][idx] = 12345
            w["edge_C"][idx] = 3.5
            tiles = plan.tiles.copy()
            tiles[pruned, 0] = -1
            w["tiles"] = torch.from_numpy(tiles).cuda()
        else:
            w["edge_D"][idx] = float("inf")
            w["edge_P"][idx] = 0
            w["edge_C"][idx] = 0.0
        for k in ("L_arr", "olap", "bits", "path", "path_len", "best_f", "best_i"):
            w[k] = torch.empty_like(ws[k])
        w["D_arr"] = torch.full_like(ws["D_arr"], -3.0)
        engine.stage2(None, plan, br.sw, w, None)
        torch.cuda.synchronize()
        outs.append(w)
    a, b = outs  # edges filled, tiles marked
    # slots of D that are defined (read by some step) after stage 2
    rd = np.ones((plan.ntiles, 2, L), bool)
    for p in range(len(pairs)):
        n1, n2, base = (int(v) for v in plan.pairs[p, 2:5])
        for i in range(n1):
            for j in range(n2):
                t = base + i * n2 + j
                if not pruned[t]:
                    continue
                R, W = int(plan.tiles[t, 2]), int(plan.tiles[t, 3])
                need0 = i + 1 >= n1 or not pruned[t + n2]
                need1 = j + 1 >= n2 or not pruned[t + 1]
                rd[t] = False
                rd[t, 0, [0, W - 1]] = True
                rd[t, 1, [0, R - 1]] = True
                rd[t, 0] |= need0
                rd[t, 1] |= need1
    rd = torch.from_numpy(rd.reshape(-1)).cuda()
    assert torch.equal(a["D_arr"][rd].view(torch.int64), b["D_arr"][rd].view(torch.int64))
    assert bool((b["D_arr"][~rd] == -3.0).all())
    for k in ("bits", "best_f", "best_i", "path_len"):
        assert torch.equal(a[k].view(torch.int64) if a[k].dtype == torch.float64 else a[k],
                           b[k].view(torch.int64) if b[k].dtype == torch.float64 else b[k]), k
    # path lengths where a path arrives (finite D): the L of an unreached slot
    # is never read
    fin = torch.isfinite(a["D_arr"]) & rd
    assert torch.equal(a["L_arr"][fin].view(torch.int64), b["L_arr"][fin].view(torch.int64))
    # (this form keeps no overlap-cost table: the sweep takes the overlap
    # costs of the unmarked tiles from edge C, and those of marked tiles are
    # 0 without a read)
    ra, rb = engine.collect(plan, a), engine.collect(plan, b)
    for p in range(len(pairs)):
        assert np.array_equal(ra.paths[p], rb.paths[p]), p


def test_listed_stage2_equals_full_stage2(torch_gpu):
    """Stage 2 with each tile diagonal launched over a list of its tiles
    (live tiles first, the others only for their corner copies), and as one
    tile-DAG launch over the live tiles, gives the outputs of the full stage
    2, bit for bit, on a batch with 40% of the tiles marked as having had no
    stage 1, and with none marked."""
    from parflexdtw_amd import engine
    from parflexdtw_amd.engine import BatchRunner
    torch = torch_gpu
    pairs = _pairs([(1500, 1700), (900, 1250), (2100, 1900), (300, 2500)], 777)
    L = 300
    br = BatchRunner(pairs, L, STEPS, WEIGHTS, beta=0.1, dtype=torch.bfloat16)
    br.launch()
    torch.cuda.synchronize()
    plan, ws = br.plan, br.ws
    rng = np.random.default_rng(11)
    for frac in (0.4, 0.0):
        pruned = rng.random(plan.ntiles) < frac
        tiles = plan.tiles.copy()
        tiles[pruned, 0] = -1
        outs = []
        for listed in (False, True, "dag"):
            w = dict(ws)
            w["tiles"] = torch.from_numpy(tiles).cuda()
            for k in ("L_arr", "olap", "bits", "path", "path_len", "best_f", "best_i",
                      "cand_D", "cand_L"):
                w[k] = torch.empty_like(ws[k])
            # the D slots of marked tiles that no step reads stay unwritten
            w["D_arr"] = torch.full_like(ws["D_arr"], -3.0)
            if listed == "dag":
                keep = engine.stage2_listed(plan, br.sw, w, ~pruned, dag=True)
            elif listed:
                keep = engine.stage2_listed(plan, br.sw, w, ~pruned)
            else:
                engine.stage2(None, plan, br.sw, w, None)
            torch.cuda.synchronize()
            engine.device_status()
            outs.append(w)
        a = outs[0]
        for b in outs[1:]:
            for k in ("D_arr", "bits", "best_f", "best_i", "path_len", "cand_D", "cand_L"):
                x, y = a[k], b[k]
                if x.dtype == torch.float64:
                    x, y = x.view(torch.int64), y.view(torch.int64)
                assert torch.equal(x, y), (frac, k)
            fin = torch.isfinite(a["D_arr"]) & (a["D_arr"] != -3.0)
            assert torch.equal(a["L_arr"][fin].view(torch.int64),
                               b["L_arr"][fin].view(torch.int64))
            ra, rb = engine.collect(plan, a), engine.collect(plan, b)
            for p in range(len(pairs)):
                assert np.array_equal(ra.paths[p], rb.paths[p]), (frac, p)


def test_batch_runner_stage2_dag_equals_diagonal_launches(torch_gpu):
    """BatchRunner(s2_dag=True): stage 2's propagation as one tile-DAG launch
    over every tile gives the per-diagonal launches' outputs bit for bit,
    eager and replayed from HIP graphs (three replays)."""
    from parflexdtw_amd.engine import BatchRunner, collect
    torch = torch_gpu
    pairs = _pairs([(1500, 1700), (900, 1250), (2100, 1900), (300, 2500)], 778)
    outs = []
    for dag in (False, True):
        br = BatchRunner(pairs, 300, STEPS, WEIGHTS, beta=0.1, dtype=torch.bfloat16, s2_dag=dag)
        br.launch()
        torch.cuda.synchronize()
        outs.append((br, collect(br.plan, br.ws)))
    (a, ra), (b, rb) = outs
    for k in ("D_arr", "best_f", "best_i", "path_len"):
        x, y = a.ws[k], b.ws[k]
        if x.dtype == torch.float64:
            x, y = x.view(torch.int64), y.view(torch.int64)
        assert torch.equal(x, y), k
    for p in range(len(pairs)):
        assert np.array_equal(ra.paths[p], rb.paths[p]), p
    b.capture()
    for _ in range(3):
        b.ws["D_arr"].fill_(-3.0)
        b.replay()
        torch.cuda.synchronize()
        assert torch.equal(a.ws["D_arr"].view(torch.int64), b.ws["D_arr"].view(torch.int64))
        rr = collect(b.plan, b.ws)
        for p in range(len(pairs)):
            assert np.array_equal(ra.paths[p], rr.paths[p]), p
