"""GPU tests of the learned tile proposal kernels (csrc/pf_learn.hip) against
the CPU oracle (oracle/tile_net_oracle.py), on identical BF16 operands.

Tolerances: pooled features, pooled cost and tile scores are computed in the
same order on both sides and must match exactly; a convolution or weight
gradient differs only by FP32 accumulation inside the matrix cores
(|diff| <= 1e-5 of the largest value); a BF16 output may differ by one BF16
rounding step; the end-to-end gradient of the 4 layers, where those
roundings feed each other, within 2% in norm."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from oracle import tile_net_oracle as tn
    from parflexdtw_amd import tile_net
    return torch, tn, tile_net


def _bf(torch, a):
    return torch.from_numpy(np.asarray(a, np.float32)).to(torch.bfloat16).cuda()


def _lib():
    from parflexdtw_amd._lib import check, lib
    return lib(), check


def test_pooled_input_matches_oracle(env):
    torch, tn, tile_net = env
    from parflexdtw_amd.synth import make_pair
    F1, F2, _ = make_pair(333, 300, K=12, seed=3)
    net = tile_net.TileNet(seed=0)
    X, H, W, P = net.pooled_input(torch.from_numpy(F1).cuda(), torch.from_numpy(F2).cuda(),
                                  want_cost=True)
    Aref, Bref = tn.pool_features(F1, 32), tn.pool_features(F2, 32)
    Pref = tn.pooled_cost(Aref, Bref)
    assert (H, W) == Pref.shape
    # block sums in another order (lane shuffles): float32 rounding apart
    assert np.allclose(P.cpu().numpy().reshape(H, W), Pref, rtol=0, atol=2e-6)
    Xh = X.float().cpu().numpy().reshape(H, W, 32)
    ref = tn.network_input(Pref)
    assert np.all(np.abs(Xh[..., 0] - ref) <= 2 ** -8 * np.abs(ref) + 1e-4)
    assert not Xh[..., 1:].any()
    Xc, _ = net.input_from_pooled(*net.pooled_features(torch.from_numpy(F1).cuda(),
                                                       torch.from_numpy(F2).cuda()),
                                  compact=True)
    assert np.array_equal(Xc.float().cpu().numpy().reshape(H, W), Xh[..., 0])


def _batch_pairs(shapes, seed):
    """Nonnegative random features (12, n) of the given ragged lengths."""
    rng = np.random.default_rng(seed)
    return [(rng.random((12, a)) ** 3, rng.random((12, b)) ** 3) for a, b in shapes]


BATCH_SHAPES = [(333, 300), (32, 65), (31, 1), (1500, 2100), (1, 40), (700, 517)]


def test_batched_pooling_and_statistics_equal_per_pair(env):
    """pf_pool_features_batch over concatenated ragged pairs equals
    pf_pool_features on each pair, and pf_input_stats_batch equals the
    statistics pf_pooled_cost computes for the pair alone, bit for bit."""
    torch, tn, tile_net = env
    L, check = _lib()
    import ctypes
    pairs = _batch_pairs(BATCH_SHAPES, 50)
    K = 12
    st = torch.cuda.current_stream().cuda_stream
    F1 = torch.from_numpy(np.concatenate([a for a, _ in pairs], axis=1)).cuda()
    F2 = torch.from_numpy(np.concatenate([b for _, b in pairs], axis=1)).cuda()
    n1 = [a.shape[1] for a, _ in pairs]
    n2 = [b.shape[1] for _, b in pairs]
    H = [-(-v // 32) for v in n1]
    W = [-(-v // 32) for v in n2]
    f1o, f2o = np.cumsum([0] + n1[:-1]), np.cumsum([0] + n2[:-1])
    hb, wb = np.cumsum([0] + H[:-1]), np.cumsum([0] + W[:-1])
    nbA, nbB = sum(H), sum(W)
    seg1 = torch.from_numpy(np.stack([f1o, n1, hb, 0 * hb], 1).astype(np.int64)).cuda()
    seg2 = torch.from_numpy(np.stack([f2o, n2, wb, 0 * wb], 1).astype(np.int64)).cuda()
    Ap = torch.empty(K * nbA, dtype=torch.float32, device="cuda")
    Bp = torch.empty(K * nbB, dtype=torch.float32, device="cuda")
    check(L.pf_pool_features_batch(F1.data_ptr(), K, F1.shape[1], seg1.data_ptr(), len(pairs), 32,
                                   Ap.data_ptr(), nbA, st), "pool A")
    check(L.pf_pool_features_batch(F2.data_ptr(), K, F2.shape[1], seg2.data_ptr(), len(pairs), 32,
                                   Bp.data_ptr(), nbB, st), "pool B")
    stp = torch.from_numpy(np.stack([hb, H, wb, W], 1).astype(np.int64)).cuda()
    stats = torch.empty(len(pairs) * (2 * K + 1), dtype=torch.float32, device="cuda")
    check(L.pf_input_stats_batch(Ap.data_ptr(), nbA, Bp.data_ptr(), nbB, K, stp.data_ptr(),
                                 len(pairs), stats.data_ptr(), st), "stats")
    Ah = Ap.cpu().numpy().reshape(K, nbA)
    Bh = Bp.cpu().numpy().reshape(K, nbB)
    sh = stats.cpu().numpy().reshape(len(pairs), 2 * K + 1)
    net = tile_net.TileNet(seed=0)
    n = ctypes.c_int64()
    check(L.pf_pooled_cost_workspace(K, ctypes.byref(n)), "ws")
    for p, (a, b) in enumerate(pairs):
        A1, B1, h, w, _ = net.pooled_features(torch.from_numpy(a).cuda(),
                                              torch.from_numpy(b).cuda())
        assert np.array_equal(Ah[:, hb[p]:hb[p] + h], A1.cpu().numpy().reshape(K, h)), p
        assert np.array_equal(Bh[:, wb[p]:wb[p] + w], B1.cpu().numpy().reshape(K, w)), p
        work = torch.empty(n.value, dtype=torch.float32, device="cuda")
        Xc = torch.empty(h * w, dtype=torch.bfloat16, device="cuda")
        check(L.pf_pooled_cost(A1.data_ptr(), h, B1.data_ptr(), w, K, None, Xc.data_ptr(), None,
                               work.data_ptr(), st), "pooled cost")
        assert np.array_equal(sh[p], work.cpu().numpy()[:2 * K + 1]), p


@pytest.mark.parametrize("K", [12, 20])
def test_pooling_with_bf16_features_equals_separate_launches(env, K):
    """pf_pool_features_bf16_batch gives the pooled features of
    pf_pool_features_batch and the BF16 features of pf_features_bf16 (two
    chunks of 16 for K = 20), bit for bit, over ragged concatenated pairs."""
    torch, tn, tile_net = env
    L, check = _lib()
    rng = np.random.default_rng(70 + K)
    ns = [333, 32, 31, 1, 1500, 65]
    F = torch.from_numpy(rng.random((K, sum(ns))) ** 3).cuda()
    nb_ = [-(-v // 32) for v in ns]
    fo, bo = np.cumsum([0] + ns[:-1]), np.cumsum([0] + nb_[:-1])
    nb = sum(nb_)
    seg = torch.from_numpy(np.stack([fo, ns, bo, 0 * bo], 1).astype(np.int64)).cuda()
    st = torch.cuda.current_stream().cuda_stream
    Kp = -(-K // 16) * 16
    n = F.shape[1]
    P0 = torch.empty(K * nb, dtype=torch.float32, device="cuda")
    P1 = torch.full((K * nb,), -1.0, dtype=torch.float32, device="cuda")
    B0 = torch.empty(n * Kp, dtype=torch.bfloat16, device="cuda")
    B1 = torch.full((n * Kp,), 7.0, dtype=torch.bfloat16, device="cuda")
    check(L.pf_pool_features_batch(F.data_ptr(), K, n, seg.data_ptr(), len(ns), 32,
                                   P0.data_ptr(), nb, st), "pool")
    check(L.pf_features_bf16(F.data_ptr(), K, n, B0.data_ptr(), st), "features")
    check(L.pf_pool_features_bf16_batch(F.data_ptr(), K, n, seg.data_ptr(), len(ns), 32,
                                        P1.data_ptr(), nb, B1.data_ptr(), st), "pool + features")
    torch.cuda.synchronize()
    assert torch.equal(P0.view(torch.int32), P1.view(torch.int32))
    assert torch.equal(B0.view(torch.int16), B1.view(torch.int16))


def test_first_layer_from_pooled_features_equals_input_map_path(env):
    """pf_conv3x3_first_pooled on a pair inside a batch equals the input map
    of pf_pooled_cost followed by pf_conv3x3 mode 3, bit for bit."""
    torch, tn, tile_net = env
    L, check = _lib()
    import ctypes
    pairs = _batch_pairs([(333, 300), (1500, 2100), (31, 1)], 60)
    K = 12
    st = torch.cuda.current_stream().cuda_stream
    net = tile_net.TileNet(seed=2)
    pooled = [net.pooled_features(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
              for a, b in pairs]
    # a batch layout with the pairs' pooled columns side by side
    H = [q[2] for q in pooled]
    W = [q[3] for q in pooled]
    lda, ldb = sum(H) + 3, sum(W) + 5
    Ap = torch.zeros((K, lda), dtype=torch.float32, device="cuda")
    Bp = torch.zeros((K, ldb), dtype=torch.float32, device="cuda")
    hb, wb = np.cumsum([0] + H[:-1]) + 2, np.cumsum([0] + W[:-1]) + 1
    for p, (A1, B1, h, w, _) in enumerate(pooled):
        Ap[:, hb[p]:hb[p] + h] = A1.reshape(K, h)
        Bp[:, wb[p]:wb[p] + w] = B1.reshape(K, w)
    n = ctypes.c_int64()
    check(L.pf_pooled_cost_workspace(K, ctypes.byref(n)), "ws")
    bias = net.layer(0)[32 * 288:]
    for p, (A1, B1, h, w, _) in enumerate(pooled):
        work = torch.empty(n.value, dtype=torch.float32, device="cuda")
        Xc = torch.empty(h * w, dtype=torch.bfloat16, device="cuda")
        check(L.pf_pooled_cost(A1.data_ptr(), h, B1.data_ptr(), w, K, None, Xc.data_ptr(), None,
                               work.data_ptr(), st), "pooled cost")
        Yref = torch.empty(h * w * 32, dtype=torch.bfloat16, device="cuda")
        check(L.pf_conv3x3(3, Xc.data_ptr(), h, w, net.Wt[0].data_ptr(), bias.data_ptr(), None,
                           Yref.data_ptr(), st), "first")
        Y = torch.empty(h * w * 32, dtype=torch.bfloat16, device="cuda")
        check(L.pf_conv3x3_first_pooled(Ap.data_ptr(), lda, int(hb[p]), h, Bp.data_ptr(), ldb,
                                        int(wb[p]), w, K, work.data_ptr(), net.Wt[0].data_ptr(),
                                        bias.data_ptr(), Y.data_ptr(), st), "first pooled")
        assert torch.equal(Y.view(torch.int16), Yref.view(torch.int16)), p


@pytest.mark.parametrize("H,W", [(13, 70), (1, 5), (40, 33)])
def test_conv_modes_match_oracle(env, H, W):
    torch, tn, _ = env
    L, check = _lib()
    rng = np.random.default_rng(H * 100 + W)
    X = tn.bf16(rng.standard_normal((H, W, 32)))
    M = tn.bf16(np.maximum(rng.standard_normal((H, W, 32)), 0))
    Wt = tn.bf16(rng.standard_normal((32, 288)) * 0.1)
    b = rng.standard_normal(32).astype(np.float32)
    Xd, Md, Wd = _bf(torch, X), _bf(torch, M), _bf(torch, Wt)
    Wfd = _bf(torch, tn.flip(Wt))
    bd = torch.from_numpy(b).cuda()
    ref = tn.conv(X, Wt, b.astype(np.float64))
    tol = 1e-5 * np.abs(ref).max() + 1e-6
    st = torch.cuda.current_stream().cuda_stream
    Z = torch.empty(H * W * 32, dtype=torch.float32, device="cuda")
    check(L.pf_conv3x3(1, Xd.data_ptr(), H, W, Wd.data_ptr(), bd.data_ptr(), None, Z.data_ptr(),
                       st), "conv1")
    Y = torch.empty(H * W * 32, dtype=torch.bfloat16, device="cuda")
    check(L.pf_conv3x3(0, Xd.data_ptr(), H, W, Wd.data_ptr(), bd.data_ptr(), None, Y.data_ptr(),
                       st), "conv0")
    G = torch.empty(H * W * 32, dtype=torch.bfloat16, device="cuda")
    check(L.pf_conv3x3(2, Xd.data_ptr(), H, W, Wfd.data_ptr(), None, Md.data_ptr(), G.data_ptr(),
                       st), "conv2")
    Zh = Z.cpu().numpy()[:H * W].reshape(H, W)        # mode 1: channel 0 only
    assert np.max(np.abs(Zh - ref[..., 0])) <= tol
    Yh = Y.float().cpu().numpy().reshape(H, W, 32)
    relu = np.maximum(ref, 0)
    assert np.all(np.abs(Yh - relu) <= 2 ** -7 * np.abs(relu) + tol)
    gref = np.where(M > 0, tn.conv(X, tn.flip(Wt)), 0.0)
    Gh = G.float().cpu().numpy().reshape(H, W, 32)
    assert np.all(np.abs(Gh - gref) <= 2 ** -7 * np.abs(gref) + 1e-5 * np.abs(gref).max() + 1e-6)


def test_layer_sweeps_equal_band_by_band(env):
    """On an image large enough for a workgroup to sweep several blocks of a
    column strip (1100 x 1560 pooled cells), every layer kernel equals the
    same kernel run band by band (bands of at most 400 rows with one halo
    row on each side, so one block per workgroup), bit for bit: the sweep
    and the prefetch of the next halo change nothing in the arithmetic."""
    torch, tn, tile_net = env
    lib_, check = _lib()
    H, W, K = 1100, 1560, 12
    rng = np.random.default_rng(77)
    st = torch.cuda.current_stream().cuda_stream
    X = torch.from_numpy(np.maximum(rng.standard_normal((H, W, 32)), 0).astype(np.float32))
    X = X.to(torch.bfloat16).cuda().contiguous()
    M = torch.from_numpy(rng.standard_normal((H, W, 32)).astype(np.float32)).to(
        torch.bfloat16).cuda().contiguous()
    net = tile_net.TileNet(seed=9)
    Wt, Wf = net.Wt[1], net.Wf[1]
    bias = net.layer(1)[32 * 288:]

    def run(mode, Xs, Ms, h):
        out = torch.empty((h, W) if mode == 1 else (h, W, 32),
                          dtype=torch.float32 if mode == 1 else torch.bfloat16, device="cuda")
        check(lib_.pf_conv3x3(mode, Xs.data_ptr(), h, W, (Wf if mode == 2 else Wt).data_ptr(),
                              None if mode == 2 else bias.data_ptr(),
                              Ms.data_ptr() if mode == 2 else None, out.data_ptr(), st), "conv")
        return out

    bands = [(a, min(a + 400, H)) for a in range(0, H, 400)]
    for mode in (0, 1, 2):
        full = run(mode, X, M, H)
        for a, b in bands:
            lo, hi = max(a - 1, 0), min(b + 1, H)
            part = run(mode, X[lo:hi].contiguous(), M[lo:hi].contiguous(), hi - lo)
            assert torch.equal(full[a:b].view(torch.int16) if mode != 1 else full[a:b],
                               part[a - lo:b - lo].view(torch.int16) if mode != 1
                               else part[a - lo:b - lo]), (mode, a)
    # first layer from pooled features
    A = torch.from_numpy(rng.random((K, H)).astype(np.float32) / 12).cuda()
    B = torch.from_numpy(rng.random((K, W)).astype(np.float32) / 12).cuda()
    stats = torch.empty(2 * K + 1, dtype=torch.float32, device="cuda")
    pairs = torch.tensor([[0, H, 0, W]], dtype=torch.int64, device="cuda")
    check(lib_.pf_input_stats_batch(A.data_ptr(), H, B.data_ptr(), W, K, pairs.data_ptr(), 1,
                                    stats.data_ptr(), st), "stats")
    b0 = net.layer(0)[32 * 288:]

    def first(a_off, h):
        out = torch.empty((h, W, 32), dtype=torch.bfloat16, device="cuda")
        check(lib_.pf_conv3x3_first_pooled(A.data_ptr(), H, a_off, h, B.data_ptr(), W, 0, W, K,
                                           stats.data_ptr(), net.Wt[0].data_ptr(), b0.data_ptr(),
                                           out.data_ptr(), st), "first")
        return out

    full = first(0, H)
    for a, b in bands:
        lo, hi = max(a - 1, 0), min(b + 1, H)
        part = first(lo, hi - lo)
        assert torch.equal(full[a:b].view(torch.int16), part[a - lo:b - lo].view(torch.int16)), a


@pytest.mark.parametrize("H,W", [(9, 31), (20, 70)])
def test_wgrad_matches_oracle(env, H, W):
    torch, tn, _ = env
    L, check = _lib()
    import ctypes
    rng = np.random.default_rng(7 + H)
    dY = tn.bf16(rng.standard_normal((H, W, 32)))
    X = tn.bf16(rng.standard_normal((H, W, 32)))
    dWr, dbr = tn.wgrad(dY, X)
    n = ctypes.c_int64()
    check(L.pf_conv3x3_wgrad_workspace(H, W, ctypes.byref(n)), "ws")
    part = torch.empty(n.value, dtype=torch.float32, device="cuda")
    g = torch.full((32 * 288 + 32,), 0.5, dtype=torch.float32, device="cuda")
    dYd, Xd = _bf(torch, dY), _bf(torch, X)   # keep both alive during the call
    check(L.pf_conv3x3_wgrad(dYd.data_ptr(), Xd.data_ptr(), H, W, g.data_ptr(), part.data_ptr(),
                             torch.cuda.current_stream().cuda_stream), "wgrad")
    torch.cuda.synchronize()
    gh = g.cpu().numpy().astype(np.float64) - 0.5     # accumulates into the buffer
    tol = 1e-5 * np.abs(dWr).max() + 1e-5
    assert np.max(np.abs(gh[:32 * 288].reshape(32, 288) - dWr)) <= tol
    assert np.max(np.abs(gh[32 * 288:] - dbr)) <= 1e-5 * np.abs(dbr).max() + 1e-5


def test_bce_and_tile_scores_match_oracle(env):
    torch, tn, _ = env
    L, check = _lib()
    rng = np.random.default_rng(11)
    H, W = 37, 45
    Z = rng.standard_normal((H, W, 32)).astype(np.float32) * 3
    lab = (rng.random((H, W)) < 0.1).astype(np.uint8)
    Zd = torch.from_numpy(np.ascontiguousarray(Z[..., 0])).cuda()   # the logit map
    labd = torch.from_numpy(lab).cuda()
    dZ = torch.empty(H * W * 32, dtype=torch.bfloat16, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    check(L.pf_bce(Zd.data_ptr(), labd.data_ptr(), H * W, 5.0, 1.0 / (H * W), dZ.data_ptr(),
                   loss.data_ptr(), st), "bce")
    lr, gr = tn.bce(Z[..., 0].astype(np.float64), lab, 5.0, 1.0 / (H * W))
    assert float(loss.item()) == pytest.approx(lr, rel=1e-5)
    dh = dZ.float().cpu().numpy().reshape(H, W, 32)
    assert np.all(np.abs(dh[..., 0] - gr) <= 2 ** -8 * np.abs(gr) + 1e-9)
    assert not dh[..., 1:].any()
    # (L1, L2) must pool to (H, W) = (37, 45)
    for (L_, L1, L2, ext) in [(100, 1180, 1440, 0), (40, 1160, 1420, 3), (2000, 1170, 1439, 8),
                              (257, 1180, 1440, 2)]:
        s = torch.empty(-(-(L1 - 1) // (L_ - 1)) * -(-(L2 - 1) // (L_ - 1)), dtype=torch.float32,
                        device="cuda")
        check(L.pf_tile_scores(Zd.data_ptr(), H, W, 32, L_, L1, L2, ext, s.data_ptr(), st),
              "scores")
        want = tn.tile_scores(Z, 32, L_, L1, L2, ext).astype(np.float32)
        assert np.array_equal(s.cpu().numpy(), want)


def test_network_gradient_matches_oracle(env):
    """Forward + backward of the 4 layers on the GPU against the oracle's
    float64 backward with BF16 rounding at the same points."""
    torch, tn, tile_net = env
    rng = np.random.default_rng(5)
    H, W = 23, 29
    net = tile_net.TileNet(seed=3)
    wh = net.weights_host().reshape(4, -1).astype(np.float64)
    # a less trivial last layer than the initial prior
    wh[3, 32 * 288] = 0.0
    net.w.copy_(torch.from_numpy(wh.reshape(-1).astype(np.float32)))
    net.pack()
    Ws = [wh[i, :32 * 288].reshape(32, 288) for i in range(4)]
    bs = [wh[i, 32 * 288:].astype(np.float32).astype(np.float64) for i in range(4)]
    X0 = np.zeros((H, W, 32))
    X0[..., 0] = tn.bf16(rng.uniform(0.1, 1.2, (H, W)))
    lab = (rng.random((H, W)) < 0.15).astype(np.uint8)
    Xd = _bf(torch, X0).reshape(-1)
    acts, Z = net.forward(Xd, H, W)
    net.backward(acts, Z, H, W, torch.from_numpy(lab).cuda(), 4.0, 1.0 / (H * W))
    torch.cuda.synchronize()
    racts, rZ = tn.forward(X0, Ws, bs)
    Zh = Z.cpu().numpy().reshape(H, W)
    assert np.max(np.abs(Zh - rZ[..., 0])) <= 2e-2 * np.abs(rZ[..., 0]).max() + 1e-4
    # inference path: compact one-channel input and the first-layer kernel
    acts_c, Zc = net.forward(Xd.reshape(H, W, 32)[..., 0].contiguous().reshape(-1), H, W)
    torch.cuda.synchronize()
    assert np.max(np.abs(Zc.cpu().numpy().reshape(H, W) - Zh)) <= 1e-2 * np.abs(Zh).max() + 1e-3
    A1 = acts_c[1].float().cpu().numpy().reshape(H, W, 32)
    A1ref = acts[1].float().cpu().numpy().reshape(H, W, 32)
    assert np.all(np.abs(A1 - A1ref) <= 2 ** -7 * np.abs(A1ref) + 1e-5)
    rloss, dWs, dbs = tn.backward(racts, rZ, Ws, lab, 4.0, 1.0 / (H * W))
    assert float(net.loss.item()) == pytest.approx(rloss, rel=1e-2)
    g = net.g.cpu().numpy().reshape(4, -1).astype(np.float64)
    for i in range(4):
        ref = np.concatenate([dWs[i].reshape(-1), dbs[i]])
        err = np.linalg.norm(g[i] - ref) / max(np.linalg.norm(ref), 1e-30)
        assert err < 2e-2, (i, err)


def test_adam_matches_numpy(env):
    torch, _, tile_net = env
    net = tile_net.TileNet(seed=1)
    w0 = net.weights_host().astype(np.float64)
    rng = np.random.default_rng(2)
    fz = tile_net.TileNet.frozen_mask().astype(bool)
    m = np.zeros_like(w0)
    v = np.zeros_like(w0)
    w = w0.copy()
    for step in (1, 2, 3):
        gr = rng.standard_normal(w0.size).astype(np.float32)
        net.g.copy_(torch.from_numpy(gr))
        net.adam(lr=1e-2)
        g = gr.astype(np.float64)
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        upd = 1e-2 * (m / (1 - 0.9 ** step)) / (np.sqrt(v / (1 - 0.999 ** step)) + 1e-8)
        w = np.where(fz, w, w - upd)
    assert np.allclose(net.weights_host(), w, rtol=1e-5, atol=1e-6)
    assert not net.g.any()


def test_tile_scores_large_map_match_oracle(env):
    """Regions of 60-90 pooled rows and columns (the 4-rows-in-flight loop,
    rows left over, a second pass along x) give the oracle's maxima."""
    torch, tn, _ = env
    L, check = _lib()
    rng = np.random.default_rng(12)
    for (L_, L1, L2, ext) in [(2000, 6390, 6300, 8), (1000, 6390, 6300, 0), (2000, 4001, 6300, 3)]:
        H, W = -(-L1 // 32), -(-L2 // 32)
        Z = rng.standard_normal((H, W)).astype(np.float32) * 3
        Zd = torch.from_numpy(Z).cuda()
        s = torch.empty(-(-(L1 - 1) // (L_ - 1)) * -(-(L2 - 1) // (L_ - 1)), dtype=torch.float32,
                        device="cuda")
        check(L.pf_tile_scores(Zd.data_ptr(), H, W, 32, L_, L1, L2, ext, s.data_ptr(),
                               torch.cuda.current_stream().cuda_stream), "scores")
        want = tn.tile_scores(Z[..., None], 32, L_, L1, L2, ext).astype(np.float32)
        assert np.array_equal(s.cpu().numpy(), want), (L_, L1, L2, ext)
