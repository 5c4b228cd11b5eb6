"""Tile and pair tables for a batch of alignment pairs sharing one tile length.

Tiling follows Parflex.run_flexdtw_on_tiles (Parflex.py:388-427): hop = L-1,
n1 = ceil((L1-1)/hop), n2 = ceil((L2-1)/hop), tile (i, j) covers rows
[i*hop, min(i*hop+L, L1)) and columns [j*hop, min(j*hop+L, L2)).  Tiles are
numbered pair by pair in i-major order (the reference's dict order).

The tables are the int64 layouts documented in include/parflex_hip.h; this
module only does host-side bookkeeping (sizes and offsets) and allocates the
device workspace as torch tensors.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

__all__ = ["BatchPlan", "tile_grid", "tile_bounds", "buffer_global", "MAX_STITCH_TILES"]

MAX_STITCH_TILES = 101  # Parflex.py:1043


def tile_grid(L1, L2, L):
    hop = L - 1
    return math.ceil((L1 - 1) / hop), math.ceil((L2 - 1) / hop)


def tile_bounds(i, j, L1, L2, L):
    hop = L - 1
    s1, s2 = i * hop, j * hop
    return s1, min(s1 + L, L1), s2, min(s2 + L, L2)


def buffer_global(L1, L2, beta):
    """Parflex.py:1526 / Parflex_gpu.py:672-674, same evaluation order."""
    return min(L1, L2) * (1 - (1 - beta) * min(L1, L2) / max(L1, L2))


@dataclass
class BatchPlan:
    shapes: list            # [(L1, L2)]
    L: int
    beta: float = 0.1
    c_offsets: list = None  # element offset of each pair's C in the C buffer
    ldcs: list = None       # row stride of each pair's C
    tiles: np.ndarray = field(init=False)
    pairs: np.ndarray = field(init=False)

    def __post_init__(self):
        L = int(self.L)
        if L < 2:
            raise ValueError("tile length L must be at least 2")
        if L > 8192:
            raise ValueError("tile length L above 8192 is not supported")
        P = len(self.shapes)
        if P == 0:
            raise ValueError("empty batch")
        if self.c_offsets is None:
            offs, o = [], 0
            for (L1, L2) in self.shapes:
                offs.append(o)
                o += L1 * L2
            self.c_offsets = offs
        if self.ldcs is None:
            self.ldcs = [L2 for (_, L2) in self.shapes]
        hop = L - 1
        blocks = []
        pairs = np.zeros((P, 16), np.int64)
        b_off = 0
        lk_off = seg_off = path_off = 0
        ntile = 0
        self.grids = []
        cells = 0
        for p, (L1, L2) in enumerate(self.shapes):
            L1, L2 = int(L1), int(L2)
            if L1 < 2 or L2 < 2:
                raise ValueError(f"pair {p}: cost matrix must be at least 2x2, got {L1}x{L2}")
            n1, n2 = tile_grid(L1, L2, L)
            self.grids.append((n1, n2))
            base = ntile
            ldc = int(self.ldcs[p])
            co = int(self.c_offsets[p])
            # i-major tile order: tile (i, j) is number base + i*n2 + j
            s1 = np.repeat(np.arange(n1, dtype=np.int64) * hop, n2)
            s2 = np.tile(np.arange(n2, dtype=np.int64) * hop, n1)
            R = np.minimum(s1 + L, L1) - s1
            W = np.minimum(s2 + L, L2) - s2
            words = ((R + 15) // 16) * W
            t = np.empty((n1 * n2, 8), np.int64)
            t[:, 0] = co + s1 * ldc + s2
            t[:, 1] = ldc
            t[:, 2] = R
            t[:, 3] = W
            t[:, 4] = b_off + np.concatenate([[0], np.cumsum(words)[:-1]])
            t[:, 5] = (base + np.arange(n1 * n2, dtype=np.int64)) * 2 * L
            t[:, 6] = p
            t[:, 7] = (s1 // hop << 32) | (s2 // hop)
            blocks.append(t)
            b_off += int(words.sum())
            cells += int((R * W).sum())
            ntile += n1 * n2
            max_segs = max(-(-L2 // L), -(-L1 // L), 1)
            buf = int(buffer_global(L1, L2, self.beta))
            pairs[p] = [L1, L2, n1, n2, base, co, ldc, buf, lk_off, seg_off, max_segs,
                        path_off, 0, 0, 0, 0]
            lk_off += L1 + L2
            seg_off += 2 * max_segs
            path_off += L1 + L2 + 8
        self.tiles = np.ascontiguousarray(np.concatenate(blocks).reshape(-1, 8))
        self.pairs = pairs
        self.ntiles = ntile
        self.npairs = P
        self.b_words = b_off
        self.lk_total = lk_off
        self.seg_total = seg_off
        self.path_total = path_off
        self.max_diag = max(n1 + n2 - 1 for n1, n2 in self.grids)
        self.max_tile_diag = max(min(n1, n2) for n1, n2 in self.grids)
        self.max_tiles_pair = max(n1 * n2 for n1, n2 in self.grids)
        self.max_segs = int(pairs[:, 10].max())
        self.cells = cells

    def tile_index(self, p, i, j):
        return int(self.pairs[p, 4]) + i * int(self.pairs[p, 3]) + j

    def allocate(self, device, b_words=None, tiles=None, pairs=None):
        """Device workspace for stage 1 and stage 2 (torch tensors); b_words
        overrides the size of the packed back-pointer buffer, and device
        tables given as tiles / pairs are used instead of uploads of the
        plan's."""
        import torch
        L, T, P = self.L, self.ntiles, self.npairs
        bitw = (L + 32) // 32
        f64, i64, i32 = torch.float64, torch.int64, torch.int32
        ws = {
            "tiles": torch.from_numpy(self.tiles).to(device) if tiles is None else tiles,
            "pairs": torch.from_numpy(self.pairs).to(device) if pairs is None else pairs,
            "Bp": torch.empty(max(self.b_words if b_words is None else b_words, 1), dtype=i32,
                              device=device),
            "edge_D": torch.empty(T * 2 * L, dtype=f64, device=device),
            "edge_P": torch.empty(T * 2 * L, dtype=i32, device=device),
            "edge_C": torch.empty(T * 2 * L, dtype=f64, device=device),
            "D_arr": torch.empty(T * 2 * L, dtype=f64, device=device),
            "L_arr": torch.empty(T * 2 * L, dtype=f64, device=device),
            "olap": torch.empty(T * 2 * L, dtype=f64, device=device),
            "bits": torch.empty(T * 2 * bitw, dtype=i32, device=device),
            "cand_D": torch.empty(self.lk_total, dtype=f64, device=device),
            "cand_L": torch.empty(self.lk_total, dtype=f64, device=device),
            "seg_f": torch.empty(self.seg_total * 4, dtype=f64, device=device),
            "seg_i": torch.empty(self.seg_total * 8, dtype=i64, device=device),
            "best_f": torch.empty(P, dtype=f64, device=device),
            "best_i": torch.empty(P * 8, dtype=i64, device=device),
            "chain": torch.empty(P * MAX_STITCH_TILES * 8, dtype=i64, device=device),
            "pts": torch.empty(P * MAX_STITCH_TILES * (2 * L + 8) * 2, dtype=i64,
                               device=device),
            "path": torch.empty(self.path_total * 2, dtype=i64, device=device),
            "path_len": torch.empty(P, dtype=i64, device=device),
        }
        return ws
