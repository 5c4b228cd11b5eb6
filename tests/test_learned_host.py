"""Host-side logic of the learned pipeline (no GPU): tile coverage of a path
with a guard band, tile selection with a margin, and the fallback rule."""
import numpy as np
import pytest


def _brute_tiles(r, c, L, L1, L2, g):
    hop = L - 1
    n1, n2 = -(-(L1 - 1) // hop), -(-(L2 - 1) // hop)
    out = set()
    for rr, cc in zip(r, c):
        for i in range(n1):
            for j in range(n2):
                r0, r1 = i * hop, min(i * hop + L, L1) - 1
                c0, c1 = j * hop, min(j * hop + L, L2) - 1
                if max(r0, rr - g) <= min(r1, rr + g) and max(c0, cc - g) <= min(c1, cc + g):
                    out.add(i * n2 + j)
    return out, n1, n2


def test_path_tiles_matches_brute_force():
    from parflexdtw_amd.learned import path_tiles
    rng = np.random.default_rng(0)
    for _ in range(150):
        L = int(rng.integers(3, 40))
        L1, L2 = int(rng.integers(2, 200)), int(rng.integers(2, 200))
        g = int(rng.integers(0, max(1, (L - 2) // 2)))
        k = int(rng.integers(1, 25))
        r, c = rng.integers(0, L1, k), rng.integers(0, L2, k)
        want, n1, n2 = _brute_tiles(r, c, L, L1, L2, g)
        got = set(path_tiles(np.stack([r, c]), L, n1, n2, g, L1, L2).tolist())
        assert got == want


def test_select_tiles_margin():
    from parflexdtw_amd.tile_net import select_tiles
    s = np.full((5, 6), -10.0)
    s[2, 3] = 1.0
    assert select_tiles(s, 5, 6, 0.0).sum() == 1
    m = select_tiles(s, 5, 6, 0.0, margin=1)
    assert m.sum() == 9 and m[1:4, 2:5].all()
    s[0, 0] = 2.0
    m = select_tiles(s, 5, 6, 0.0, margin=1)
    assert m[0:2, 0:2].all() and m.sum() == 13


def test_fallback_rule():
    from parflexdtw_amd.learned import needs_fallback
    L, L1, L2 = 101, 401, 401     # hop 100: 4 x 4 tiles, rows [0,101) [100,201) [200,301) ...
    # a diagonal path passes the tile corners (100,100), (200,200), (300,300):
    # each corner cell lies in four tiles
    diag = np.stack([np.arange(401), np.arange(401)])
    band = np.abs(np.arange(4)[:, None] - np.arange(4)[None, :]) <= 1
    assert needs_fallback(diag, True, 0.1, band, L, 0.2, 0, L1, L2) is None
    assert needs_fallback(diag, True, 0.1, np.eye(4, dtype=bool), L, 0.2, 0, L1, L2) == \
        "selection edge"
    # a path along row 130 lies in tile row 1 only; a 40-frame guard reaches row 90 (tile row 0)
    row = np.stack([np.full(401, 130), np.arange(401)])
    sel = np.zeros((4, 4), bool)
    sel[1] = True
    assert needs_fallback(row, True, 0.1, sel, L, 0.2, 0, L1, L2) is None
    assert needs_fallback(row, True, 0.1, sel, L, 0.2, 40, L1, L2) == "selection edge"
    sel[0] = True
    assert needs_fallback(row, True, 0.1, sel, L, 0.2, 40, L1, L2) is None
    assert needs_fallback(row, True, 0.3, sel, L, 0.2, 0, L1, L2) == "cost bound"
    assert needs_fallback(row, False, np.nan, sel, L, 0.2, 0, L1, L2) == "no endpoint"


def test_pruned_near_path_matches_tile_set():
    from parflexdtw_amd.learned import _pruned_near_path, path_tiles
    rng = np.random.default_rng(3)
    for _ in range(300):
        L = int(rng.integers(3, 30))
        L1, L2 = int(rng.integers(2, 150)), int(rng.integers(2, 150))
        hop = L - 1
        n1, n2 = -(-(L1 - 1) // hop), -(-(L2 - 1) // hop)
        g = int(rng.integers(0, max(1, (L - 2) // 2 - 8)))
        # a monotone path with steps of at most 2 frames, as FlexDTW paths are
        k = int(rng.integers(1, 60))
        wp = np.stack([np.minimum(np.cumsum(rng.integers(1, 3, k)) - 1, L1 - 1),
                       np.minimum(np.cumsum(rng.integers(0, 3, k)), L2 - 1)])
        sel = rng.random((n1, n2)) < 0.8
        # the subsampled test is conservative: it flags at least every pruned
        # tile within g frames, and nothing beyond g + 8 frames
        near = not sel.reshape(-1)[path_tiles(wp, L, n1, n2, g, L1, L2)].all()
        far = not sel.reshape(-1)[path_tiles(wp, L, n1, n2, g + 8, L1, L2)].all()
        got = _pruned_near_path(wp, sel, L, g, L1, L2)
        assert (not near or got) and (far or not got)
