"""GPU parity: the HIP path against the reference's golden vectors and the CPU
oracle.  The plain form is FP64 and must match bit for bit (integers, paths
and float64 values identical).  Every test calls the product through its
public Python drop-ins, which reach libparflex_hip.so through the C ABI."""
import numpy as np
import pytest

from conftest import STAGE2_CASES, load_golden

pytestmark = pytest.mark.gpu

STEPS = [(1, 1), (1, 2), (2, 1)]


def _same(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.dtype.kind == "f" or b.dtype.kind == "f":
        return a.shape == b.shape and np.array_equal(a.astype(np.float64), b.astype(np.float64),
                                                     equal_nan=True)
    return a.shape == b.shape and np.array_equal(a, b)


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from parflexdtw_amd import Parflex_gpu, gpu_flexdtw
    return gpu_flexdtw, Parflex_gpu


def test_tile_reference_seeds(gpu):
    """tests/test_gpu_flexdtw.py:20-39 of the reference, made exact."""
    G, _ = gpu
    g = load_golden("flexdtw_seed1")
    for k in range(3):
        C = g[f"C{k}"]
        C_dev = G.cost_matrix_to_gpu(C)
        bc, wp, D, P, B, debug = G.flexdtw_chunk_from_global_C(
            C_dev, C.shape[1], 0, C.shape[0], 0, C.shape[1], STEPS, [2, 3, 3], buffer=1)
        assert _same(D, g[f"D{k}"])
        assert _same(P, g[f"P{k}"])
        assert _same(B, g[f"B{k}"])
        assert _same(wp, g[f"wp{k}"])
        assert bc == g[f"best{k}"]
        assert _same(np.array([list(d) for d in debug], np.float64), g[f"debug{k}"])
    for k in range(3):
        C = g[f"xC{k}"]
        bc, wp, D, P, B, _ = G.flexdtw_chunk_from_global_C(
            G.cost_matrix_to_gpu(C), C.shape[1], 0, C.shape[0], 0, C.shape[1], g[f"xsteps{k}"],
            g[f"xweights{k}"], buffer=1)
        assert _same(D, g[f"xD{k}"]) and _same(P, g[f"xP{k}"]) and _same(B, g[f"xB{k}"])
        assert _same(wp, g[f"xwp{k}"]) and bc == g[f"xbest{k}"]


@pytest.mark.parametrize("R,W,seed", [(3, 2, 0), (2, 300, 1), (700, 2, 2), (129, 131, 3),
                                      (257, 2048, 4), (64, 2049, 5), (300, 4096, 6),
                                      (40, 4097, 7), (33, 8192, 8), (1, 17, 9)])
def test_tile_shapes_vs_oracle(gpu, oracle_mod, R, W, seed):
    """Every lane-width variant (2/4/8 columns per lane), ragged widths, one-row
    and two-column tiles, on a sub-block of a larger matrix (ldc != W)."""
    G, _ = gpu
    rng = np.random.RandomState(seed)
    big = rng.rand(R + 5, W + 7)
    r0, c0 = 3, 5
    bc, wp, D, P, B, debug = G.flexdtw_chunk_from_global_C(
        G.cost_matrix_to_gpu(big), big.shape[1], r0, r0 + R, c0, c0 + W, STEPS, [2, 3, 3],
        buffer=1)
    bo, wpo, Do, Po, Bo, dbo = oracle_mod.flexdtw(big[r0:r0 + R, c0:c0 + W], STEPS, [2, 3, 3], 1)
    assert _same(D, Do) and _same(P, Po) and _same(B, Bo)
    if R > 1:
        assert bc == bo and _same(wp, wpo)
        assert [d[1] for d in debug] == [d[1] for d in dbo] or np.array_equal(
            np.array([d[1] for d in debug]), np.array([d[1] for d in dbo]), equal_nan=True)


def test_run_flexdtw_on_tiles_matches_reference(gpu):
    """tests/test_gpu_flexdtw.py:42-62 of the reference, made exact."""
    _, PG = gpu
    g = load_golden("tiles_seed2")
    chunks, L, n1, n2 = PG.run_flexdtw_on_tiles(g["C"], L=14, steps=STEPS, weights=[2, 3, 3],
                                                profile_dir=None)
    assert (n1, n2) == (int(g["n1"]), int(g["n2"]))
    for (i, j), ch in chunks.items():
        assert _same(ch["D"], g[f"tile_{i}_{j}_D"])
        assert _same(ch["S"], g[f"tile_{i}_{j}_S"])
        assert _same(ch["B"], g[f"tile_{i}_{j}_B"])
        assert ch["best_cost"] == g[f"tile_{i}_{j}_best"]
        assert _same(ch["wp"], g[f"tile_{i}_{j}_wp"])


def test_gather_matches_dense_reads(gpu):
    """tests/test_gpu_flexdtw.py:65-89: device gather equals host C reads."""
    G, _ = gpu
    g = load_golden("edges_seed4")
    C = g["C"]
    rng = np.random.RandomState(0)
    rows = rng.randint(0, C.shape[0], 200)
    cols = rng.randint(0, C.shape[1], 200)
    vals = G.gather_c_at_global_indices(G.cost_matrix_to_gpu(C), rows, cols)
    assert _same(vals, C[rows, cols])


@pytest.mark.parametrize("name", STAGE2_CASES)
def test_parflex_matches_reference(gpu, name):
    _, PG = gpu
    g = load_golden(name)
    C, L, beta = g["C"], int(g["L"]), float(g["beta"])
    if not int(g["found"]):
        with pytest.raises(ValueError, match="No valid endpoint"):
            PG.parflex(C, g["steps"], g["weights"], beta, L, profile_dir=None)
        return
    bc, wp = PG.parflex(C, g["steps"], g["weights"], beta, L, profile_dir=None)
    assert bc == g["parflex_best_cost"]
    assert _same(wp, g["parflex_wp"])


@pytest.mark.parametrize("name", [n for n in STAGE2_CASES if n != "stage2_nocand"])
def test_stage2_dict_matches_reference(gpu, name):
    """run_stage2_from_tiled: D_arr, L_arr, edge summary, per-segment winners
    and their stitched paths equal Parflex.run_stage2_from_tiled."""
    _, PG = gpu
    g = load_golden(name)
    C, L, beta = g["C"], int(g["L"]), float(g["beta"])
    tiled = {"C_dev": None, "C": C, "L_block": L,
             "stage1_params": {"steps": g["steps"], "weights": g["weights"], "buffer": 1.0}}
    r = PG.run_stage2_from_tiled(tiled, C, beta=beta, backtrace_segments=True)
    assert _same(r["D_arr"], g["s2_D_arr"])
    assert _same(r["L_arr"], g["s2_L_arr"])
    assert _same(r["edge_lens"], g["s2_edge_lens"])
    assert r["best_cost"] == g["s2_best_cost"]
    assert _same(np.array(r["best_end"]), g["s2_best_end"])
    assert _same(r["stitched_wp"], g["s2_stitched_wp"])
    es = r["edge_summary"]
    assert _same(es["top"]["D"], g["s2_top_D"]) and _same(es["top"]["L"], g["s2_top_L"])
    assert _same(es["right"]["D"], g["s2_right_D"]) and _same(es["right"]["L"], g["s2_right_L"])
    segs = []
    norms = []
    paths = []
    for key in sorted(r["best_per_segment"]):
        m = r["best_per_segment"][key]
        segs.append([0 if key[0] == "top" else 1, key[1], m["chunk_i"], m["chunk_j"], m["edge"],
                     m["idx"], m["global_coord"][0], m["global_coord"][1]])
        norms.append(m["norm_cost"])
        paths.append(np.array(r["paths_per_segment"][key]["path"], np.int64).reshape(-1, 2))
    assert _same(np.array(segs, np.int64).reshape(-1, 8), g["s2_segments"])
    assert _same(np.array(norms), g["s2_segment_norm"])
    assert [len(p) for p in paths] == list(g["s2_segment_path_lens"])
    if paths:
        assert _same(np.concatenate(paths), g["s2_segment_paths"])


def test_cfg0_workload_matches_reference(gpu):
    """2000 x 2400 chroma pair, L = 500: the reference's CPU result, exactly,
    with the cost matrix built on the GPU from the features."""
    G, PG = gpu
    g = load_golden("cfg0_chroma2000_L500")
    from parflexdtw_amd.synth import cost_matrix_det
    C = cost_matrix_det(g["F1"], g["F2"])
    C_dev = G.cost_matrix_from_features_gpu(g["F1"], g["F2"])
    assert _same(C_dev.cpu().numpy(), C)
    bc, wp = PG.parflex(C, g["steps"], g["weights"], float(g["beta"]), int(g["L"]),
                        profile_dir=None)
    assert bc == g["parflex_best_cost"]
    assert _same(wp, g["parflex_wp"])
    _, tiled = PG.tiled_stage1_from_features(g["F1"], g["F2"], steps=g["steps"],
                                             weights=g["weights"], beta=0.1, L=int(g["L"]))
    r = PG.run_stage2_from_tiled(tiled, None, beta=float(g["beta"]))
    assert r["best_cost"] == g["parflex_best_cost"]
    assert _same(r["stitched_wp"].T, g["parflex_wp"])


S1_CONFIGS = [  # (shape, L, steps, weights, (strips, waves, rows per exchange block))
    ((333, 410), 70, STEPS, [2, 3, 3], None),
    ((333, 410), 70, STEPS, [2, 3, 3], (1, 1, 16)),
    ((333, 410), 300, STEPS, [2, 3, 3], (3, 1, 7)),
    ((333, 410), 300, STEPS, [2, 3, 3], (2, 2, 1)),
    ((700, 1500), 700, STEPS, [1.25, 3, 3], (6, 1, 64)),
    ((700, 1500), 700, STEPS, [2, 3, 3], (4, 2, 16)),
    ((130, 3000), 2999, STEPS, [2, 3, 3], (5, 5, 16)),
    ((500, 520), 260, [(1, 1), (1, 2), (2, 1), (2, 2)], [2, 3, 3, 4], (3, 1, 5)),
    ((500, 520), 260, [(1, 0), (1, 1), (2, 1)], [1, 2, 3], (2, 2, 3)),
    ((300, 1100), 1100, STEPS, [2, 3, 3], (1, 8, 32)),
    ((300, 2100), 2048, STEPS, [2, 3, 3], (1, 16, 16)),
    ((200, 4100), 4000, STEPS, [2, 3, 3], (2, 16, 32)),
    # last tile row of 34 rows: its consumer strips stage every row at once
    ((333, 410), 300, STEPS, [2, 3, 3], (2, 1, 64)),
    # exchange block of 33 rows (not a power of two), 4-wave strips
    ((900, 700), 450, STEPS, [2, 3, 3], (2, 2, 33)),
    # 8 columns per lane: one strip of 4 waves, and two strips of 2 waves
    ((300, 2100), 2048, STEPS, [2, 3, 3], (1, 4, 32)),
    ((500, 1500), 1200, STEPS, [2, 3, 3], (2, 2, 24)),
]


@pytest.mark.parametrize("case", range(len(S1_CONFIGS)))
def test_stage1_packed_outputs_vs_oracle(gpu, oracle_mod, case):
    """Edge D/P and every packed back-pointer of a multi-tile stage 1, for
    one-workgroup tiles and for tiles cut into pipelined column strips
    (ragged last strips, idle strips, exchange blocks of 1..64 rows, generic
    and specialised step sets)."""
    import torch
    from parflexdtw_amd import engine
    from parflexdtw_amd._plan import BatchPlan
    shape, L, steps, weights, cfg = S1_CONFIGS[case]
    rng = np.random.RandomState(21 + case)
    C = rng.rand(*shape)
    plan = BatchPlan([C.shape], L)
    sw = engine.StepWeights(steps, weights)
    C_dev = torch.from_numpy(C).cuda()
    ws = plan.allocate(C_dev.device)
    engine.stage1(C_dev, plan, sw, ws, config=cfg)
    s1 = oracle_mod.run_stage1(C, L, steps, weights)
    n1, n2 = plan.grids[0]
    eD = ws["edge_D"].cpu().numpy().reshape(n1, n2, 2, L)
    eP = ws["edge_P"].cpu().numpy().reshape(n1, n2, 2, L)
    eC = ws["edge_C"].cpu().numpy().reshape(n1, n2, 2, L)
    Bp = ws["Bp"].cpu().numpy().view(np.uint32)
    for i in range(n1):
        for j in range(n2):
            t = plan.tiles[i * n2 + j]
            R, W, boff = int(t[2]), int(t[3]), int(t[4])
            r0, c0 = i * (L - 1), j * (L - 1)
            # first row and first column of the tile's C (start cells)
            assert _same(eC[i, j, 0, :W], C[r0, c0:c0 + W])
            assert _same(eC[i, j, 1, :R], C[r0:r0 + R, c0])
            assert _same(eD[i, j, 0, :W], s1["edge_D"][i, j, 0, :W])
            assert _same(eD[i, j, 1, :R], s1["edge_D"][i, j, 1, :R])
            assert _same(eP[i, j, 0, :W], s1["edge_P"][i, j, 0, :W])
            assert _same(eP[i, j, 1, :R], s1["edge_P"][i, j, 1, :R])
            words = Bp[boff:boff + ((R + 15) // 16) * W].reshape(-1, W)
            r = np.arange(R)
            codes = (words[r // 16, :] >> (2 * (r % 16))[:, None].astype(np.uint32)) & 3
            ob = s1["B"][i, j].reshape(-1)[:R * W].reshape(R, W)  # tile stored R x W
            assert _same(codes.astype(np.int8), ob)


FP32_CONFIGS = [  # (launch shape, L, C shape)
    (None, 200, (333, 410)), ((1, 1, 16), 200, (333, 410)), ((2, 1, 16), 200, (333, 410)),
    ((2, 1, 7), 150, (333, 410)), ((1, 2, 32), 90, (333, 410)),
    # the learned form's launch shape (one strip of 16 waves, 2 columns per
    # lane, 64-row exchange blocks) on 2048-column tiles of at most 150 rows
    ((1, 16, 64), 2048, (150, 2100)),
    # 4 columns per lane (one strip of 4 waves) and 8 (one strip of 2 waves)
    ((1, 4, 16), 1000, (180, 1100)), ((1, 2, 16), 1000, (180, 1100)),
    # two strips of 8 waves with exchange blocks of 64 rows
    ((2, 8, 64), 2048, (190, 2100)),
]


@pytest.mark.parametrize("case", range(len(FP32_CONFIGS)))
@pytest.mark.parametrize("dtype", ["float32", "bfloat16"])
def test_stage1_fp32_integer_costs_vs_oracle(gpu, oracle_mod, case, dtype):
    """The FP32 DP (standard steps) on integer costs 0..7 and tiles of at
    most 200 rows: every sum and cross product is exact in FP32 and the FP64
    quotients are distinct wherever they differ, so edges and every packed
    back-pointer must equal the FP64 oracle's (ragged last rows and columns,
    one and two strips, 2, 4 and 8 columns per lane, up to 16 waves)."""
    import torch
    from parflexdtw_amd import engine
    from parflexdtw_amd._plan import BatchPlan
    cfg, L, shape = FP32_CONFIGS[case]
    rng = np.random.RandomState(77 + case)
    C = rng.randint(0, 8, size=shape).astype(np.float64)
    plan = BatchPlan([C.shape], L)
    sw = engine.StepWeights(STEPS, [2, 3, 3])
    C_dev = torch.from_numpy(C).to(getattr(torch, dtype)).cuda()
    ws = plan.allocate(C_dev.device)
    engine.stage1(C_dev, plan, sw, ws, config=cfg)
    s1 = oracle_mod.run_stage1(C, L, STEPS, [2, 3, 3])
    n1, n2 = plan.grids[0]
    eD = ws["edge_D"].cpu().numpy().reshape(n1, n2, 2, L)
    eP = ws["edge_P"].cpu().numpy().reshape(n1, n2, 2, L)
    Bp = ws["Bp"].cpu().numpy().view(np.uint32)
    for i in range(n1):
        for j in range(n2):
            t = plan.tiles[i * n2 + j]
            R, W, boff = int(t[2]), int(t[3]), int(t[4])
            assert _same(eD[i, j, 0, :W], s1["edge_D"][i, j, 0, :W]), (i, j)
            assert _same(eD[i, j, 1, :R], s1["edge_D"][i, j, 1, :R]), (i, j)
            assert _same(eP[i, j, 0, :W], s1["edge_P"][i, j, 0, :W]), (i, j)
            assert _same(eP[i, j, 1, :R], s1["edge_P"][i, j, 1, :R]), (i, j)
            words = Bp[boff:boff + ((R + 15) // 16) * W].reshape(-1, W)
            r = np.arange(R)
            codes = (words[r // 16, :] >> (2 * (r % 16))[:, None].astype(np.uint32)) & 3
            ob = s1["B"][i, j].reshape(-1)[:R * W].reshape(R, W)
            assert _same(codes.astype(np.int8), ob), (i, j)


def test_batched_pairs_match_single(gpu, oracle_mod):
    """Several ragged pairs in one stage-1 launch and one stage-2 pass."""
    import torch
    from parflexdtw_amd import engine
    from parflexdtw_amd._plan import BatchPlan
    from parflexdtw_amd.synth import cost_matrix_det, make_pair
    Cs = []
    for p, (n, m) in enumerate([(180, 200), (260, 230), (90, 111), (301, 333)]):
        F1, F2, _ = make_pair(n, m, seed=300 + p)
        Cs.append(cost_matrix_det(F1, F2))
    flat = np.concatenate([c.reshape(-1) for c in Cs])
    plan = BatchPlan([c.shape for c in Cs], 40, beta=0.1)
    res, _ = engine.run_batch(torch.from_numpy(flat).cuda(), plan, STEPS, [2, 3, 3])
    for p, C in enumerate(Cs):
        bc, wp = oracle_mod.parflex(C, STEPS, [2, 3, 3], 0.1, 40)
        assert res.found[p]
        assert res.best_cost[p] == bc
        assert _same(res.paths[p], wp)


@pytest.mark.parametrize("cfg", ["cfg1"])
def test_full_size_single_pair_vs_oracle(gpu, oracle_mod, cfg):
    """16000 x ~19000 chroma pair, L = 4000 (25 tiles, ~3e8 cells): the GPU
    path equals the oracle index for index; best cost identical."""
    G, PG = gpu
    from parflexdtw_amd.synth import cfg_pair
    F1, F2, _ = cfg_pair(1)
    C_dev = G.cost_matrix_from_features_gpu(F1, F2)
    C = C_dev.cpu().numpy()
    bc, wp = PG.parflex(C_dev, STEPS, [2, 3, 3], 0.1, 4000, profile_dir=None)
    bo, wpo = oracle_mod.parflex(C, STEPS, [2, 3, 3], 0.1, 4000, nthreads=16)
    assert bc == bo
    assert _same(wp, wpo)


def _bf16_round(x):
    """Round float64 values to BF16 (nearest even) through float32."""
    f = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    f = (f + 0x7FFF + ((f >> 16) & 1)) & 0xFFFF0000
    return f.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("K,N,M", [(12, 300, 170), (128, 257, 129), (20, 64, 64), (3, 33, 200)])
def test_cost_mfma_bf16_layout_and_error(gpu, K, N, M):
    """MFMA cost build: FP32 output equals the product of the BF16-rounded
    normalised features to FP32 accumulation accuracy (a wrong MFMA operand or
    accumulator layout gives O(1) errors); BF16 output within BF16 rounding of
    the FP64 cost."""
    import torch
    G, _ = gpu
    from parflexdtw_amd.synth import cost_matrix_det, make_pair
    F1, F2, _ = make_pair(N, M + 20, K=K, seed=K + N, kind="relu")
    F2 = np.ascontiguousarray(F2[:, :M])
    n1 = np.sqrt(np.sum(F1 * F1, axis=0)) + 1e-9
    n2 = np.sqrt(np.sum(F2 * F2, axis=0)) + 1e-9
    A = _bf16_round(F1 / n1)
    B = _bf16_round(F2 / n2)
    ref_bf = 1.0 - A.T @ B
    C32 = G.cost_matrix_from_features_gpu(F1, F2, dtype=torch.float32).cpu().numpy()
    assert np.max(np.abs(C32 - ref_bf)) < 2e-5
    C16 = G.cost_matrix_from_features_gpu(F1, F2, dtype=torch.bfloat16).float().cpu().numpy()
    ref64 = cost_matrix_det(F1, F2)
    assert np.max(np.abs(C16 - ref64)) < 1.6e-2
    assert np.max(np.abs(C32 - ref64)) < 1e-2


@pytest.mark.parametrize("extra", [0, 1, 19642])
def test_one_tile_tile_length_at_least_matrix(gpu, oracle_mod, extra):
    """L >= max(L1, L2): one tile covers the matrix (Parflex.py:388-391),
    also for L above the multi-tile limit; the result, and the stage-2 dict
    with its (1, 1, 2, L) arrays, equal the oracle's at the caller's L."""
    _, PG = gpu
    from parflexdtw_amd.synth import cost_matrix_det, make_pair
    F1, F2, _ = make_pair(300, 380, seed=17)
    C = cost_matrix_det(F1, F2)
    L = max(C.shape) + extra
    bc, wp = PG.parflex(C, STEPS, [2, 3, 3], 0.1, L, profile_dir=None)
    bo, wpo = oracle_mod.parflex(C, STEPS, [2, 3, 3], 0.1, L)
    assert bc == bo and _same(wp, wpo)
    tiled = {"C_dev": None, "C": C, "L_block": L,
             "stage1_params": {"steps": STEPS, "weights": [2, 3, 3], "buffer": 1.0}}
    r = PG.run_stage2_from_tiled(tiled, C, beta=0.1)
    s1 = oracle_mod.run_stage1(C, L, STEPS, [2, 3, 3])
    ro = oracle_mod.run_stage2(C, s1, 0.1)
    assert r["D_arr"].shape == (1, 1, 2, L)
    assert _same(r["D_arr"], ro["D_arr"]) and _same(r["L_arr"], ro["L_arr"])
    assert _same(r["stitched_wp"], ro["stitched_wp"])
