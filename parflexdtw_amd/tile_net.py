"""Learned tile proposal: a 4-layer 3x3 convolution network (32 channels)
over the 1/32-pooled cost matrix that scores every tile by how likely the
alignment path crosses it; only tiles above a threshold, plus a margin, run
stage 1.  The reference computes every tile, so this module has no reference
counterpart; the plain full-tile method is its fallback and the reference it is checked against.

Every kernel is in csrc/pf_learn.hip: pooled features and pooled cost, the
convolution layers on MFMA (BF16 operands, FP32 accumulation), the weight
gradient, the weighted BCE loss, Adam on FP32 master weights, and the
per-tile scores.  torch only holds the buffers.

Layout of the parameters (f32 master copy, gradient and Adam moments): four
layers of [32 cout][288] weights (k = (dy*3+dx)*32 + cin) followed by 32
biases.  The input carries the pooled cost in channel 0; layer 1's weights on
the other input channels and layer 4's outputs other than channel 0 are
frozen at zero.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from ._lib import check, lib

__all__ = ["TileNet", "POOL", "path_block_labels", "select_tiles"]

POOL = 32
CH = 32
KDIM = 9 * CH
NPER = CH * KDIM + CH      # parameters of one layer (weights + biases)
NLAYERS = 4


def _torch():
    import torch
    return torch


def _st(stream):
    torch = _torch()
    return (stream or torch.cuda.current_stream()).cuda_stream


def path_block_labels(wp, L1, L2, pool=POOL):
    """uint8 (ceil(L1/pool), ceil(L2/pool)): 1 where the path (2, n) crosses
    the pool x pool block."""
    H, W = -(-L1 // pool), -(-L2 // pool)
    lab = np.zeros((H, W), np.uint8)
    wp = np.asarray(wp)
    if wp.size:
        lab[wp[0] // pool, wp[1] // pool] = 1
    return lab


def select_tiles(scores, n1, n2, threshold, margin=0):
    """Boolean (n1, n2): score >= threshold, dilated by `margin` whole tiles
    in every direction."""
    s = np.asarray(scores).reshape(n1, n2) >= threshold
    if margin <= 0 or not s.any():
        return s
    out = s.copy()
    for di in range(-margin, margin + 1):
        for dj in range(-margin, margin + 1):
            src = s[max(0, -di):n1 - max(0, di), max(0, -dj):n2 - max(0, dj)]
            out[max(0, di):max(0, di) + src.shape[0], max(0, dj):max(0, dj) + src.shape[1]] |= src
    return out


class TileNet:
    def __init__(self, device=None, seed=0, weights=None):
        torch = _torch()
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        f32 = torch.float32
        if weights is None:
            weights = self.init_weights(seed)
        w = np.ascontiguousarray(weights, np.float32).reshape(-1)
        if w.size != NLAYERS * NPER:
            raise ValueError(f"expected {NLAYERS * NPER} parameters, got {w.size}")
        self.w = torch.from_numpy(w.copy()).to(self.device)
        self.g = torch.zeros_like(self.w)
        self.m = torch.zeros_like(self.w)
        self.v = torch.zeros_like(self.w)
        self.frozen = torch.from_numpy(self.frozen_mask()).to(self.device)
        self.Wt = [torch.empty(CH * KDIM, dtype=torch.bfloat16, device=self.device)
                   for _ in range(NLAYERS)]
        self.Wf = [torch.empty(CH * KDIM, dtype=torch.bfloat16, device=self.device)
                   for _ in range(NLAYERS)]
        self.step = 0
        self._part = None
        self.loss = torch.zeros(1, dtype=f32, device=self.device)
        self.pack()

    # ---- parameters -------------------------------------------------
    @staticmethod
    def frozen_mask():
        fz = np.zeros((NLAYERS, NPER), np.uint8)
        w1 = fz[0, :CH * KDIM].reshape(CH, 9, CH)
        w1[:, :, 1:] = 1                       # layer 1 sees only input channel 0
        w4 = fz[3, :CH * KDIM].reshape(CH, KDIM)
        w4[1:, :] = 1                          # layer 4: only output channel 0 is used
        fz[3, CH * KDIM + 1:] = 1
        return fz.reshape(-1)

    @staticmethod
    def init_weights(seed=0):
        rng = np.random.default_rng(seed)
        w = np.zeros((NLAYERS, NPER), np.float32)
        w1 = w[0, :CH * KDIM].reshape(CH, 9, CH)
        w1[:, :, 0] = rng.standard_normal((CH, 9)) * math.sqrt(2.0 / 9)
        for layer in (1, 2):
            w[layer, :CH * KDIM] = rng.standard_normal(CH * KDIM) * math.sqrt(2.0 / KDIM)
        w[3, :KDIM] = rng.standard_normal(KDIM) * math.sqrt(1.0 / KDIM)
        w[3, CH * KDIM] = -3.0                 # prior: few blocks are on the path
        return w.reshape(-1)

    def layer(self, i):
        return self.w[i * NPER:(i + 1) * NPER]

    def pack(self, stream=None):
        """BF16 weights (and their flipped copies) from the master weights."""
        for i in range(NLAYERS):
            check(lib().pf_pack_weights(self.layer(i).data_ptr(), self.Wt[i].data_ptr(),
                                        self.Wf[i].data_ptr(), _st(stream)), "pf_pack_weights")

    def weights_host(self):
        return self.w.detach().cpu().numpy().copy()

    def save(self, path, **meta):
        np.savez(path, weights=self.weights_host(),
                 **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path, device=None):
        with np.load(path, allow_pickle=False) as z:
            net = cls(device=device, weights=z["weights"])
            meta = {k: z[k] for k in z.files if k != "weights"}
        return net, meta

    # ---- input --------------------------------------------------------
    def pooled_features(self, F1d, F2d, stream=None):
        """Pooled normalised features (K*H and K*W f32) of one pair of device
        features (K, N) / (K, M) float64."""
        torch = _torch()
        K, N = int(F1d.shape[0]), int(F1d.shape[1])
        M = int(F2d.shape[1])
        H, W = -(-N // POOL), -(-M // POOL)
        A = torch.empty(K * H, dtype=torch.float32, device=self.device)
        B = torch.empty(K * W, dtype=torch.float32, device=self.device)
        st = _st(stream)
        check(lib().pf_pool_features(F1d.data_ptr(), K, N, POOL, A.data_ptr(), st),
              "pf_pool_features")
        check(lib().pf_pool_features(F2d.data_ptr(), K, M, POOL, B.data_ptr(), st),
              "pf_pool_features")
        return A, B, H, W, K

    def input_from_pooled(self, A, B, H, W, K, stream=None, want_cost=False, compact=False):
        """Network input from pooled features: H*W*32 bf16 (the layout the
        training passes use) or, with compact=True, the one-channel H*W bf16
        map the inference pass starts from; also the f32 pooled cost (H*W)
        if asked."""
        torch = _torch()
        X = torch.empty(H * W * (1 if compact else CH), dtype=torch.bfloat16, device=self.device)
        n = ctypes.c_int64()
        check(lib().pf_pooled_cost_workspace(K, ctypes.byref(n)), "pf_pooled_cost_workspace")
        work = torch.empty(n.value, dtype=torch.float32, device=self.device)
        P = torch.empty(H * W, dtype=torch.float32, device=self.device) if want_cost else None
        check(lib().pf_pooled_cost(A.data_ptr(), H, B.data_ptr(), W, K,
                                   None if compact else X.data_ptr(),
                                   X.data_ptr() if compact else None,
                                   P.data_ptr() if P is not None else None, work.data_ptr(),
                                   _st(stream)), "pf_pooled_cost")
        return X, P

    def pooled_input(self, F1d, F2d, stream=None, want_cost=False, compact=False):
        """Network input for one pair of device features (K, N) / (K, M)
        float64 (see input_from_pooled); also the f32 pooled cost if asked."""
        A, B, H, W, K = self.pooled_features(F1d, F2d, stream)
        X, P = self.input_from_pooled(A, B, H, W, K, stream, want_cost, compact)
        return (X, H, W, P) if want_cost else (X, H, W)

    # ---- forward / backward -------------------------------------------
    def forward(self, X, H, W, stream=None):
        """Activations [X, A1, A2, A3] (bf16) and the logit map Z (f32 H*W).
        X is the 32-channel input, or the one-channel compact input (then the
        first layer reads channel 0 only)."""
        torch = _torch()
        st = _st(stream)
        acts = [X]
        A = X
        first = 3 if X.numel() == H * W else 0
        for i in range(3):
            Y = torch.empty(H * W * CH, dtype=torch.bfloat16, device=self.device)
            check(lib().pf_conv3x3(first if i == 0 else 0, A.data_ptr(), H, W,
                                   self.Wt[i].data_ptr(), self.layer(i)[CH * KDIM:].data_ptr(),
                                   None, Y.data_ptr(), st), "pf_conv3x3")
            acts.append(Y)
            A = Y
        Z = torch.empty(H * W, dtype=torch.float32, device=self.device)
        check(lib().pf_conv3x3(1, A.data_ptr(), H, W, self.Wt[3].data_ptr(),
                               self.layer(3)[CH * KDIM:].data_ptr(), None, Z.data_ptr(), st),
              "pf_conv3x3")
        return acts, Z

    def backward(self, acts, Z, H, W, label, pos_weight, scale, stream=None):
        """Accumulate the loss and the parameter gradient of one image into
        self.loss and self.g."""
        torch = _torch()
        st = _st(stream)
        npix = H * W
        need = ctypes.c_int64()
        check(lib().pf_conv3x3_wgrad_workspace(H, W, ctypes.byref(need)),
              "pf_conv3x3_wgrad_workspace")
        if self._part is None or self._part.numel() < need.value:
            self._part = torch.empty(need.value, dtype=torch.float32, device=self.device)
        dY = torch.empty(npix * CH, dtype=torch.bfloat16, device=self.device)
        check(lib().pf_bce(Z.data_ptr(), label.data_ptr(), npix, float(pos_weight), float(scale),
                           dY.data_ptr(), self.loss.data_ptr(), st), "pf_bce")
        for i in (3, 2, 1, 0):
            check(lib().pf_conv3x3_wgrad(dY.data_ptr(), acts[i].data_ptr(), H, W,
                                         self.g[i * NPER:(i + 1) * NPER].data_ptr(),
                                         self._part.data_ptr(), st), "pf_conv3x3_wgrad")
            if i == 0:
                break
            dX = torch.empty(npix * CH, dtype=torch.bfloat16, device=self.device)
            check(lib().pf_conv3x3(2, dY.data_ptr(), H, W, self.Wf[i].data_ptr(), None,
                                   acts[i].data_ptr(), dX.data_ptr(), st), "pf_conv3x3")
            dY = dX

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, stream=None):
        self.step += 1
        check(lib().pf_adam(self.w.data_ptr(), self.g.data_ptr(), self.m.data_ptr(),
                            self.v.data_ptr(), self.frozen.data_ptr(), self.w.numel(), float(lr),
                            float(b1), float(b2), float(eps), self.step, _st(stream)), "pf_adam")
        self.pack(stream)

    def forward_pooled(self, Ap, lda, a_off, H, Bp, ldb, b_off, W, K, st, bufs, stream=None):
        """Logit map of one pair of a batch, from the batch's pooled features
        (Ap: K x lda, Bp: K x ldb f32; the pair's columns start at a_off,
        b_off) and its statistics (device pointer `st` to 2K+1 floats, as
        written by pf_input_stats_batch).  The first layer builds its input
        on the fly.  bufs = (X1, X2, Z): two bf16 activation buffers of at
        least H*W*32 elements and an f32 buffer of at least H*W; returns Z."""
        X1, X2, Z = bufs
        sp = _st(stream)
        check(lib().pf_conv3x3_first_pooled(Ap.data_ptr(), lda, a_off, H, Bp.data_ptr(), ldb,
                                            b_off, W, K, st, self.Wt[0].data_ptr(),
                                            self.layer(0)[CH * KDIM:].data_ptr(), X1.data_ptr(),
                                            sp), "pf_conv3x3_first_pooled")
        for i, (src, dst) in ((1, (X1, X2)), (2, (X2, X1))):
            check(lib().pf_conv3x3(0, src.data_ptr(), H, W, self.Wt[i].data_ptr(),
                                   self.layer(i)[CH * KDIM:].data_ptr(), None, dst.data_ptr(), sp),
                  "pf_conv3x3")
        check(lib().pf_conv3x3(1, X1.data_ptr(), H, W, self.Wt[3].data_ptr(),
                               self.layer(3)[CH * KDIM:].data_ptr(), None, Z.data_ptr(), sp),
              "pf_conv3x3")
        return Z

    # ---- inference -------------------------------------------------------
    def tile_scores(self, Z, H, W, L, L1, L2, ext=0, stream=None):
        """Max logit over each tile's pooled cells extended by ext cells."""
        torch = _torch()
        hop = L - 1
        n1, n2 = -(-(L1 - 1) // hop), -(-(L2 - 1) // hop)
        s = torch.empty(n1 * n2, dtype=torch.float32, device=self.device)
        check(lib().pf_tile_scores(Z.data_ptr(), H, W, POOL, L, L1, L2, int(ext), s.data_ptr(),
                                   _st(stream)), "pf_tile_scores")
        return s, n1, n2

    def score_pair(self, F1d, F2d, L, ext=0, stream=None):
        """Tile logits (n1*n2 device tensor) of one pair of device features."""
        X, H, W = self.pooled_input(F1d, F2d, stream, compact=True)
        _, Z = self.forward(X, H, W, stream)
        return self.tile_scores(Z, H, W, L, int(F1d.shape[1]), int(F2d.shape[1]), ext, stream)
