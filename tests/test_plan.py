"""Host bookkeeping of the tile/pair tables (no device work)."""
import numpy as np
import pytest

from parflexdtw_amd._plan import BatchPlan, buffer_global, tile_bounds, tile_grid


@pytest.mark.parametrize("L1,L2,L", [(40, 44, 14), (2000, 2380, 500), (16000, 19000, 4000),
                                     (10, 10, 10), (11, 30, 3), (2, 2, 2), (5, 900, 64)])
def test_tile_table_geometry(oracle_mod, L1, L2, L):
    n1, n2 = tile_grid(L1, L2, L)
    assert (n1, n2) == oracle_mod.tile_grid(L1, L2, L)
    plan = BatchPlan([(L1, L2)], L)
    t = plan.tiles
    assert t.shape == (n1 * n2, 8)
    cells = 0
    for i in range(n1):
        for j in range(n2):
            row = t[i * n2 + j]
            s1, e1, s2, e2 = tile_bounds(i, j, L1, L2, L)
            assert (s1, e1, s2, e2) == oracle_mod.tile_bounds(i, j, L1, L2, L)
            assert row[0] == s1 * L2 + s2 and row[1] == L2
            assert (row[2], row[3]) == (e1 - s1, e2 - s2)
            assert row[2] >= 2 and row[3] >= 2          # edges are never shorter than 2
            assert row[5] == (i * n2 + j) * 2 * L
            assert row[7] == (i << 32) | j
            cells += (e1 - s1) * (e2 - s2)
    assert plan.cells == cells
    # packed back-pointer words are disjoint and cover ceil(R/16)*W words each
    ends = t[:, 4] + (t[:, 2] + 15) // 16 * t[:, 3]
    assert np.all(t[1:, 4] == ends[:-1]) and plan.b_words == ends[-1]


def test_buffer_global_truncation():
    # Parflex.py:1526 then int() at Parflex.py:1248
    assert int(buffer_global(40, 44, 0.1)) == int(40 * (1 - (1 - 0.1) * 40 / 44))
    assert int(buffer_global(30, 30, 1.0)) == 30
    plan = BatchPlan([(40, 44), (100, 80)], 14, beta=0.1)
    assert plan.pairs[0, 7] == int(buffer_global(40, 44, 0.1))
    assert plan.pairs[1, 7] == int(buffer_global(100, 80, 0.1))
    assert plan.pairs[1, 4] == plan.grids[0][0] * plan.grids[0][1]


def test_plan_rejects_degenerate():
    with pytest.raises(ValueError):
        BatchPlan([(1, 10)], 5)
    with pytest.raises(ValueError):
        BatchPlan([(10, 10)], 1)
    with pytest.raises(ValueError):
        BatchPlan([], 5)


def test_stage1_tuner_candidates():
    """The features form adds one strip of 4 columns per lane (k_sweep4,
    K <= 128); every candidate covers the tile."""
    from parflexdtw_amd.engine import _config_candidates
    for L in (232, 500, 1000, 2000, 4000):
        for feat, Kp in ((False, 16), (True, 16), (True, 128), (True, 144)):
            cands = _config_candidates(L, feat, Kp)
            assert cands
            for S, nw in cands:
                assert S * nw * 64 * 8 >= L
            one4 = [(S, nw) for S, nw in cands if S == 1 and nw * 128 < L]
            assert bool(one4) == (feat and Kp <= 128), (L, feat, Kp, cands)
