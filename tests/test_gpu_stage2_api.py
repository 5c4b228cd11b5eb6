"""The reference's stage-2 functions called one by one, as a caller that
drives stage 2 step by step does (Parflex.py:862-1356, Parflex_gpu.py:247-339):
propagate_tile_edge_costs, sync_tile_overlap_edges, stage2_scan_and_stitch and
backtrace_and_stitch of parflexdtw_amd.Parflex_gpu, each on the device, against
the reference's outputs (golden vectors) bit for bit."""
import numpy as np
import pytest

from conftest import STAGE2_CASES, load_golden

pytestmark = pytest.mark.gpu


def _same(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.dtype.kind == "f" or b.dtype.kind == "f":
        return a.shape == b.shape and np.array_equal(a.astype(np.float64), b.astype(np.float64),
                                                     equal_nan=True)
    return a.shape == b.shape and np.array_equal(a, b)


@pytest.fixture(scope="module")
def PG():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from parflexdtw_amd import Parflex_gpu
    return Parflex_gpu


@pytest.fixture(scope="module")
def presync():
    return load_golden("stage2_presync")


def _check_dict(r, g):
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
    for key in sorted(r["best_per_segment"]):
        m = r["best_per_segment"][key]
        segs.append([0 if key[0] == "top" else 1, key[1], m["chunk_i"], m["chunk_j"], m["edge"],
                     m["idx"], m["global_coord"][0], m["global_coord"][1]])
    assert _same(np.array(segs, np.int64).reshape(-1, 8), g["s2_segments"])


def _step_by_step(PG, chunks, C, L, n1, n2, g, pre, name, C_dev=None):
    """The reference's stage-2 sequence (Parflex.py:1520-1541) through the
    drop-in functions."""
    tiled = PG.build_tiled_metadata(chunks, L, n1, n2, C,
                                    stage1_params={"steps": g["steps"], "weights": g["weights"],
                                                   "buffer": 1.0})
    D_arr, L_arr, edge_lens = PG.propagate_tile_edge_costs(chunks, L, n1, n2, buffer_param=1,
                                                           profile_dir=None, C_dev=C_dev)
    assert _same(D_arr, pre[f"{name}__D_arr"])
    assert _same(L_arr, pre[f"{name}__L_arr"])
    assert _same(edge_lens, pre[f"{name}__edge_lens"])
    D2, L2s = PG.sync_tile_overlap_edges(D_arr, L_arr, edge_lens, n1, n2)
    assert D2 is D_arr and L2s is L_arr
    L1, L2 = C.shape
    beta = float(g["beta"])
    buf = min(L1, L2) * (1 - (1 - beta) * min(L1, L2) / max(L1, L2))
    if not int(g["found"]):
        with pytest.raises(ValueError, match="No valid endpoint"):
            PG.stage2_scan_and_stitch(tiled, chunks, D_arr, L_arr, edge_lens, L1, L2, L_block=L,
                                      buffer_stage2=buf, profile_dir=None)
        return
    r = PG.stage2_scan_and_stitch(tiled, chunks, D_arr, L_arr, edge_lens, L1, L2, L_block=L,
                                  buffer_stage2=buf, profile_dir=None, backtrace_segments=True)
    assert r["D_arr"] is D_arr
    _check_dict(r, g)
    # backtrace_and_stitch from every segment winner (sorted, as the fixture)
    paths = []
    for key in sorted(r["best_per_segment"]):
        m = r["best_per_segment"][key]
        p = PG.backtrace_and_stitch(tiled, chunks, m["chunk_i"], m["chunk_j"], m["edge"],
                                    m["idx"])
        assert p == r["paths_per_segment"][key]["path"]
        paths.append(np.array(p, np.int64).reshape(-1, 2))
    assert [len(p) for p in paths] == list(g["s2_segment_path_lens"])
    if paths:
        assert _same(np.concatenate(paths), g["s2_segment_paths"])
    wp = r["stitched_wp"].T
    assert _same(wp, g["parflex_wp"]) and r["best_cost"] == g["parflex_best_cost"]


@pytest.mark.parametrize("name", STAGE2_CASES)
def test_stage2_step_by_step_on_host_chunks(PG, presync, name):
    """chunks_dict of run_flexdtw_on_tiles (host D, S, B per tile)."""
    g = load_golden(name)
    C, L = g["C"], int(g["L"])
    chunks, L_out, n1, n2 = PG.run_flexdtw_on_tiles(C, L, steps=g["steps"], weights=g["weights"],
                                                    profile_dir=None)
    _step_by_step(PG, chunks, C, L_out, n1, n2, g, presync, name)


def test_stage2_step_by_step_on_reference_chunks(PG, presync):
    """A chunks_dict holding the reference's own per-tile D, S, B (the
    fixture of run_flexdtw_on_tiles, tests/test_gpu_flexdtw.py:42-62) and no
    C: the overlap costs come from the tiled_result's C."""
    from parflexdtw_amd._plan import tile_bounds
    g = load_golden("tiles_seed2")
    C, L = g["C"], int(g["L"])
    n1, n2 = int(g["n1"]), int(g["n2"])
    chunks = {}
    for i in range(n1):
        for j in range(n2):
            s1, e1, s2, e2 = tile_bounds(i, j, C.shape[0], C.shape[1], L)
            chunks[(i, j)] = {"C": None, "D": g[f"tile_{i}_{j}_D"], "S": g[f"tile_{i}_{j}_S"],
                              "B": g[f"tile_{i}_{j}_B"], "wp": g[f"tile_{i}_{j}_wp"],
                              "best_cost": float(g[f"tile_{i}_{j}_best"]),
                              "bounds": (s1, e1, s2, e2), "shape": (e1 - s1, e2 - s2)}
    # propagate has no C argument besides C_dev: hand it the device matrix
    import torch
    C_dev = torch.from_numpy(C).cuda()
    D_arr, L_arr, edge_lens = PG.propagate_tile_edge_costs(chunks, L, n1, n2, profile_dir=None,
                                                           C_dev=C_dev)
    assert _same(D_arr, presync["tiles_seed2__D_arr"])
    assert _same(L_arr, presync["tiles_seed2__L_arr"])
    _step_by_step(PG, chunks, C, L, n1, n2, g, presync, "tiles_seed2", C_dev=C_dev)


@pytest.mark.parametrize("name", ["stage2_chroma300_L50", "stage2_chroma420_L37",
                                  "stage2_chroma600_L100", "stage2_chroma400_L5_cap"])
def test_stage2_step_by_step_on_device_chunks(PG, presync, name):
    """chunks_dict of tiled_stage1_from_features: stage 1 stays on the device
    and the stage-2 functions use it directly."""
    g = load_golden(name)
    C, L = g["C"], int(g["L"])
    _, tiled = PG.tiled_stage1_from_features(g["F1"], g["F2"], steps=g["steps"],
                                             weights=g["weights"], beta=float(g["beta"]), L=L)
    chunks = tiled["chunks_dict"]
    n1, n2 = tiled["n_row"], tiled["n_col"]
    D_arr, L_arr, edge_lens = PG.propagate_tile_edge_costs(chunks, L, n1, n2, profile_dir=None)
    assert _same(D_arr, presync[f"{name}__D_arr"])
    assert _same(L_arr, presync[f"{name}__L_arr"])
    PG.sync_tile_overlap_edges(D_arr, L_arr, edge_lens, n1, n2)
    L1, L2 = C.shape
    beta = float(g["beta"])
    buf = min(L1, L2) * (1 - (1 - beta) * min(L1, L2) / max(L1, L2))
    r = PG.stage2_scan_and_stitch(tiled, chunks, D_arr, L_arr, edge_lens, L1, L2, L_block=L,
                                  buffer_stage2=buf, profile_dir=None)
    _check_dict(r, g)
    be = r["best_end"]
    p = PG.backtrace_and_stitch(tiled, chunks, be[2], be[3], be[4], be[5])
    assert _same(np.array(p, np.int64).reshape(-1, 2), g["s2_stitched_wp"])


def test_sync_on_random_arrays(PG):
    """sync_tile_overlap_edges where every pass acts (infinite slots on one
    side of an overlap, corners 0, 5e-11 and 2e-10 apart): the reference's
    outputs exactly, written back in place."""
    g = load_golden("stage2_sync_random")
    for k in range(int(g["ncases"])):
        D = g[f"c{k}_D_in"].copy()
        La = g[f"c{k}_L_in"].copy()
        L1, L2, L, n1, n2 = (int(v) for v in g[f"c{k}_shape"])
        D2, L2s = PG.sync_tile_overlap_edges(D, La, g[f"c{k}_edge_lens"], n1, n2)
        assert D2 is D and L2s is La
        assert _same(D, g[f"c{k}_D_out"]), k
        assert _same(La, g[f"c{k}_L_out"]), k


def test_backtrace_and_stitch_bad_start(PG):
    g = load_golden("tiles_seed2")
    C, L = g["C"], int(g["L"])
    chunks, L_out, n1, n2 = PG.run_flexdtw_on_tiles(C, L, profile_dir=None)
    tiled = PG.build_tiled_metadata(chunks, L_out, n1, n2, C)
    assert PG.backtrace_and_stitch(tiled, chunks, n1, 0, 0, 0) == []
    with pytest.raises(IndexError):
        PG.backtrace_and_stitch(tiled, chunks, 0, 0, 0, 10 ** 6)
