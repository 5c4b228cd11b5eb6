"""CPU oracle for the learned tile proposal -- TEST INFRASTRUCTURE ONLY.

NumPy restatement of the network the HIP kernels in
parflexdtw_amd/csrc/pf_learn.hip compute (the reference has no learned part;
this restates this project's own specification):

  pooled features  Fp[k, b] = mean over the frames i of block b of
                   F[k, i] / (sqrt(sum_k F[k, i]^2) + 1e-9)
  pooled cost      P = 1 - A^T B        (the block mean of C)
  network input    P double-centred and divided by its standard deviation
  layer            Y[p, o] = sum_{dy, dx, c} X[p + (dy-1, dx-1), c] W[o, (dy*3+dx)*32 + c] + b[o]
                   (zero padding, 32 channels NHWC), ReLU on hidden layers
  loss             weighted BCE on the channel-0 logits
  backward         dX = conv(dY, flipped W) masked by the ReLU, dW = dY^T im2col(X)

Arithmetic is float64.  Where the GPU rounds a tensor to BF16 (activations,
gradients, weights) the functions take a `bf16` flag and round at the same
place, so GPU and oracle see identical operands; the remaining difference is
FP32 accumulation inside the matrix cores.  Used only by tests/.
"""
from __future__ import annotations

import numpy as np

CH = 32
KDIM = 9 * CH


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def pool_features(F, pool):
    """Float32 block means of the normalised frames, in the GPU's order."""
    F = np.asarray(F, np.float64)
    K, n = F.shape
    nb = -(-n // pool)
    out = np.zeros((K, nb), np.float32)
    for b in range(nb):
        i0, i1 = b * pool, min(b * pool + pool, n)
        for i in range(i0, i1):
            s = 0.0
            for k in range(K):
                s = s + F[k, i] * F[k, i]
            inv = 1.0 / ((np.sqrt(s) + 1e-9) * float(i1 - i0))
            for k in range(K):
                out[k, b] = np.float32(out[k, b] + np.float32(F[k, i] * inv))
    return out


def pooled_cost(A, B):
    """1 - A^T B in float32 with the k order of the GPU (no fused multiply-add)."""
    A = np.asarray(A, np.float32)
    B = np.asarray(B, np.float32)
    s = np.zeros((A.shape[1], B.shape[1]), np.float32)
    for k in range(A.shape[0]):
        s = (s + (A[k][:, None] * B[k][None, :]).astype(np.float32)).astype(np.float32)
    return (np.float32(1.0) - s).astype(np.float32)


def network_input(P):
    """Channel 0 of the network input: P double-centred and divided by its
    standard deviation (float64; the GPU computes it in float32)."""
    P = np.asarray(P, np.float64)
    c = P - P.mean(axis=1, keepdims=True) - P.mean(axis=0, keepdims=True) + P.mean()
    return c / np.sqrt(np.mean(c * c) + 1e-12)


def im2col(X):
    """X (H, W, 32) -> (H*W, 288), column (dy*3+dx)*32 + c, zero padding."""
    H, W, C = X.shape
    Xp = np.zeros((H + 2, W + 2, C), X.dtype)
    Xp[1:-1, 1:-1] = X
    cols = [Xp[dy:dy + H, dx:dx + W, :] for dy in range(3) for dx in range(3)]
    return np.concatenate(cols, axis=2).reshape(H * W, 9 * C)


def conv(X, Wt, b=None):
    """Pre-activation of one layer: (H, W, 32) float64."""
    H, W, _ = X.shape
    Y = im2col(X) @ Wt.T
    if b is not None:
        Y = Y + b[None, :]
    return Y.reshape(H, W, CH)


def flip(Wt):
    """Wf[c, (8-dydx)*32 + o] = W[o, dydx*32 + c] (weights of the input gradient)."""
    W4 = Wt.reshape(CH, 9, CH)              # [o, dydx, c]
    return W4[:, ::-1, :].transpose(2, 1, 0).reshape(CH, KDIM)


def wgrad(dY, X):
    """dW (32, 288) and db (32,) of one layer."""
    H, W, _ = X.shape
    G = dY.reshape(H * W, CH)
    return G.T @ im2col(X), G.sum(axis=0)


def bce(z, label, pos_weight, scale):
    """Loss and logit gradient of the weighted BCE (scaled)."""
    y = label.astype(np.float64)
    w = np.where(label > 0, pos_weight, 1.0)
    loss = np.sum(w * (np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))) * scale
    g = w * (1.0 / (1.0 + np.exp(-z)) - y) * scale
    return loss, g


def forward(X0, Ws, bs, round_bf16=True):
    """Activations [X0, A1, A2, A3] and logits Z (H, W, 32) of the 4 layers."""
    r = bf16 if round_bf16 else (lambda a: a)
    acts = [X0]
    A = X0
    for layer in range(3):
        A = r(np.maximum(conv(A, r(Ws[layer]), bs[layer]), 0.0))
        acts.append(A)
    Z = conv(A, r(Ws[3]), bs[3])
    return acts, Z


def backward(acts, Z, Ws, label, pos_weight, scale, round_bf16=True):
    """Loss and gradients (dW list, db list) of forward(); the logit gradient
    and the hidden-layer gradients are rounded to BF16 where the GPU stores them."""
    r = bf16 if round_bf16 else (lambda a: a)
    loss, g = bce(Z[..., 0], label, pos_weight, scale)
    dY = np.zeros_like(Z)
    dY[..., 0] = g
    dY = r(dY)
    dWs, dbs = [None] * 4, [None] * 4
    for layer in (3, 2, 1, 0):
        dWs[layer], dbs[layer] = wgrad(dY, acts[layer])
        if layer == 0:
            break
        dX = conv(dY, flip(r(Ws[layer])))
        dY = r(np.where(acts[layer] > 0, dX, 0.0))
    return loss, dWs, dbs


def tile_scores(Z, pool, L, L1, L2, ext=0):
    """Max channel-0 logit over the pooled cells each tile covers, extended
    by ext cells on every side (i-major)."""
    hop = L - 1
    n1, n2 = -(-(L1 - 1) // hop), -(-(L2 - 1) // hop)
    H, W = Z.shape[:2]
    out = np.empty(n1 * n2)
    for i in range(n1):
        r0, r1 = i * hop, min(i * hop + L, L1) - 1
        for j in range(n2):
            c0, c1 = j * hop, min(j * hop + L, L2) - 1
            out[i * n2 + j] = Z[max(r0 // pool - ext, 0):min(r1 // pool + ext, H - 1) + 1,
                                max(c0 // pool - ext, 0):min(c1 // pool + ext, W - 1) + 1, 0].max()
    return out
