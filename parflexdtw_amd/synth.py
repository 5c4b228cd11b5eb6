"""Seeded synthetic alignment workloads (feature pairs with a known time warp).

Chroma pairs: F1 is K x N with i.i.d. Gamma(shape 0.5, scale 1) entries, each
column L2-normalised.  F2 is F1 resampled along a random monotone
piecewise-linear time warp (8 knots, per-segment tempo ratio U[0.8, 1.25]) to
M0 frames, plus N(0, 0.05^2) noise, clipped at 0, renormalised, and trimmed by
a random 0-10 % of frames at each end.  Dense pairs ("relu") use Gaussian
entries passed through ReLU instead of Gamma draws.

Features are feature-major (K, frames), the layout of the reference's
``FlexDTW.L2norm`` (FlexDTW.py:1207-1210).
"""
from __future__ import annotations

import numpy as np

__all__ = ["make_pair", "cfg_pair", "batch_pairs", "l2_normalise_columns",
           "path_agreement", "cost_matrix_det"]

STEPS = np.array([[1, 1], [1, 2], [2, 1]], dtype=np.int64)
WEIGHTS = np.array([2.0, 3.0, 3.0], dtype=np.float64)
BETA = 0.1


def l2_normalise_columns(F):
    n = np.sqrt(np.sum(F * F, axis=0))
    n = np.where(n > 0, n, 1.0)
    return F / n[None, :]


def cost_matrix_det(F1, F2):
    """C = 1 - L2norm(F1)^T L2norm(F2) (FlexDTW.py:1207-1210, Parflex.py:1470)
    with a fixed, machine-independent summation order: squares and products
    are summed over the feature index k in increasing order, each operation
    rounded on its own.  The HIP FP64 cost kernel follows the same order."""
    F1 = np.asarray(F1, dtype=np.float64)
    F2 = np.asarray(F2, dtype=np.float64)
    s1 = np.zeros(F1.shape[1])
    s2 = np.zeros(F2.shape[1])
    for k in range(F1.shape[0]):
        s1 = s1 + F1[k] * F1[k]
    for k in range(F2.shape[0]):
        s2 = s2 + F2[k] * F2[k]
    n1 = np.sqrt(s1) + 1e-9
    n2 = np.sqrt(s2) + 1e-9
    acc = np.zeros((F1.shape[1], F2.shape[1]))
    for k in range(F1.shape[0]):
        acc = acc + (F1[k] / n1)[:, None] * (F2[k] / n2)[None, :]
    return 1.0 - acc


def make_pair(N, M0, K=12, seed=0, kind="chroma", n_knots=8,
              tempo=(0.8, 1.25), noise=0.05, trim=0.10):
    """Return (F1 (K,N), F2 (K,M), warp) where warp[m] is the fractional F1
    frame that F2 frame m was sampled at."""
    rng = np.random.default_rng(seed)
    if kind == "chroma":
        F1 = rng.gamma(0.5, 1.0, size=(K, N))
    elif kind == "relu":
        F1 = np.maximum(rng.standard_normal((K, N)), 0.0)
    else:
        raise ValueError(f"unknown feature kind {kind!r}")
    F1 = l2_normalise_columns(F1)

    # Monotone piecewise-linear warp from F2 time [0, M0-1] onto F1 time
    # [0, N-1]: n_knots knots evenly spaced in F2 time, segment slopes scaled
    # by tempo ratios, then rescaled so the end points match.
    knots2 = np.linspace(0.0, M0 - 1.0, n_knots)
    ratios = rng.uniform(tempo[0], tempo[1], size=n_knots - 1)
    seg = np.diff(knots2) * ratios
    knots1 = np.concatenate([[0.0], np.cumsum(seg)])
    knots1 *= (N - 1.0) / knots1[-1]
    t2 = np.arange(M0, dtype=np.float64)
    warp = np.interp(t2, knots2, knots1)

    lo = np.floor(warp).astype(np.int64)
    lo = np.clip(lo, 0, N - 1)
    hi = np.clip(lo + 1, 0, N - 1)
    fr = warp - lo
    F2 = F1[:, lo] * (1.0 - fr)[None, :] + F1[:, hi] * fr[None, :]
    F2 = F2 + rng.normal(0.0, noise, size=F2.shape)
    F2 = np.maximum(F2, 0.0)
    F2 = l2_normalise_columns(F2)

    a = int(np.floor(rng.uniform(0.0, trim) * M0))
    b = int(np.floor(rng.uniform(0.0, trim) * M0))
    F2 = np.ascontiguousarray(F2[:, a:M0 - b])
    warp = warp[a:M0 - b]
    return np.ascontiguousarray(F1), F2, warp


def cfg_pair(cfg, p=0):
    """Feature pair for a named workload.

    cfg 0: chroma 2000 x 2400 (seed 0); cfg 1: chroma 16000 x 19000 (seed 1);
    cfg 2: chroma batch pair p, N ~ U[3000,20000], M = N*U[0.8,1.25] (seed 2000+p);
    cfg 3: relu K=128, 40000 x 40000 (seed 3).
    """
    if cfg == 0:
        return make_pair(2000, 2400, seed=0)
    if cfg == 1:
        return make_pair(16000, 19000, seed=1)
    if cfg == 2:
        rng = np.random.default_rng(2000 + p)
        N = int(rng.integers(3000, 20001))
        M = int(round(N * rng.uniform(0.8, 1.25)))
        return make_pair(N, M, seed=2000 + p)
    if cfg == 3:
        return make_pair(40000, 40000, K=128, seed=3, kind="relu")
    raise ValueError(f"unknown workload {cfg}")


def batch_pairs(n_pairs, seed_base=2000, n_lo=3000, n_hi=20000):
    out = []
    for p in range(n_pairs):
        rng = np.random.default_rng(seed_base + p)
        N = int(rng.integers(n_lo, n_hi + 1))
        M = int(round(N * rng.uniform(0.8, 1.25)))
        out.append(make_pair(N, M, seed=seed_base + p))
    return out


def path_agreement(wp_ref, wp, tol=2):
    """Fraction of reference-path F1 frames whose interpolated F2 frame on the
    candidate path lies within +-tol frames (np.interp as in eval_tools.py:45)."""
    wp_ref = np.asarray(wp_ref)
    wp = np.asarray(wp)
    if wp_ref.shape[1] == 0 or wp.shape[1] == 0:
        return 0.0
    pred = np.interp(wp_ref[0].astype(np.float64), wp[0].astype(np.float64),
                     wp[1].astype(np.float64))
    return float(np.mean(np.abs(pred - wp_ref[1]) <= tol))


def beat_times(warp, hop, every=22, offset=0):
    """Ground-truth beat pairs of a synthetic pair in seconds: reference
    (F2) frames offset, offset+every, ... and the fractional F1 frame each
    was sampled at (make_pair's warp), i.e. (query times, reference times)
    as the .beat annotations of a real pair give them."""
    m = np.arange(offset, len(warp), every)
    return np.asarray(warp, dtype=np.float64)[m] * hop, m.astype(np.float64) * hop
