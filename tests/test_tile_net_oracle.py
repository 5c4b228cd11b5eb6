"""The learned-proposal oracle (oracle/tile_net_oracle.py) pinned by textbook
routines and by the mathematics: the convolution against
scipy.signal.correlate2d, the backward pass against central finite
differences of the float64 forward, the pooled cost against the block mean of
the full cost matrix, the flip against its defining identity."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def tn():
    from oracle import tile_net_oracle
    return tile_net_oracle


def test_conv_matches_correlate2d(tn):
    from scipy.signal import correlate2d
    rng = np.random.default_rng(0)
    X = rng.standard_normal((7, 9, 32))
    Wt = rng.standard_normal((32, 288))
    b = rng.standard_normal(32)
    Y = tn.conv(X, Wt, b)
    K = Wt.reshape(32, 3, 3, 32)  # [o, dy, dx, c]
    for o in (0, 5, 31):
        want = b[o] + sum(correlate2d(X[:, :, c], K[o, :, :, c], mode="same") for c in range(32))
        assert np.allclose(Y[:, :, o], want, rtol=1e-12, atol=1e-10)


def test_flip_is_the_input_gradient_of_conv(tn):
    """<conv(X, W), G> == <X, conv(G, flip(W))> for every X, G."""
    rng = np.random.default_rng(1)
    X = rng.standard_normal((6, 5, 32))
    G = rng.standard_normal((6, 5, 32))
    Wt = rng.standard_normal((32, 288))
    lhs = np.sum(tn.conv(X, Wt) * G)
    rhs = np.sum(X * tn.conv(G, tn.flip(Wt)))
    assert lhs == pytest.approx(rhs, rel=1e-12)


def test_backward_matches_finite_differences(tn):
    rng = np.random.default_rng(2)
    H, W = 5, 6
    X0 = np.zeros((H, W, 32))
    X0[..., 0] = rng.uniform(0.2, 1.2, (H, W))
    Ws = [rng.standard_normal((32, 288)) * 0.3 for _ in range(4)]
    bs = [rng.standard_normal(32) * 0.1 + 0.2 for _ in range(4)]
    label = (rng.random((H, W)) < 0.3).astype(np.uint8)

    def loss_of(Ws_, bs_):
        acts, Z = tn.forward(X0, Ws_, bs_, round_bf16=False)
        return tn.bce(Z[..., 0], label, 4.0, 1.0 / (H * W))[0]

    acts, Z = tn.forward(X0, Ws, bs, round_bf16=False)
    _, dWs, dbs = tn.backward(acts, Z, Ws, label, 4.0, 1.0 / (H * W), round_bf16=False)
    eps = 1e-6
    for layer, (o, k) in [(0, (3, 4 * 32)), (1, (7, 100)), (2, (31, 287)), (3, (0, 17))]:
        Wp = [w.copy() for w in Ws]
        Wm = [w.copy() for w in Ws]
        Wp[layer][o, k] += eps
        Wm[layer][o, k] -= eps
        fd = (loss_of(Wp, bs) - loss_of(Wm, bs)) / (2 * eps)
        assert dWs[layer][o, k] == pytest.approx(fd, rel=1e-5, abs=1e-9)
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[layer][o] += eps
        bm[layer][o] -= eps
        fd = (loss_of(Ws, bp) - loss_of(Ws, bm)) / (2 * eps)
        assert dbs[layer][o] == pytest.approx(fd, rel=1e-5, abs=1e-9)


def test_pooled_cost_is_block_mean_of_cost(tn):
    from parflexdtw_amd.synth import cost_matrix_det, make_pair
    F1, F2, _ = make_pair(200, 170, seed=5)
    C = cost_matrix_det(F1, F2)
    P = tn.pooled_cost(tn.pool_features(F1, 32), tn.pool_features(F2, 32))
    H, W = P.shape
    for y in range(H):
        for x in range(W):
            blk = C[y * 32:(y + 1) * 32, x * 32:(x + 1) * 32]
            assert P[y, x] == pytest.approx(blk.mean(), abs=2e-6)


def test_tile_scores_geometry(tn):
    Z = np.zeros((8, 10, 32))
    Z[..., 0] = np.arange(80).reshape(8, 10)
    # L1 = 250, L2 = 300, L = 100: rows [0,100) [99,199) [198,250) ...
    s = tn.tile_scores(Z, 32, 100, 250, 300)
    hop = 99
    n1, n2 = -(-249 // hop), -(-299 // hop)
    assert s.shape == (n1 * n2,)
    for i in range(n1):
        for j in range(n2):
            r1 = min(i * hop + 100, 250) - 1
            c1 = min(j * hop + 100, 300) - 1
            assert s[i * n2 + j] == min(r1 // 32, 7) * 10 + min(c1 // 32, 9)
    # extended by 2 cells: the region's bottom-right cell moves 2 further
    s2 = tn.tile_scores(Z, 32, 100, 250, 300, ext=2)
    for i in range(n1):
        for j in range(n2):
            r1 = min(i * hop + 100, 250) - 1
            c1 = min(j * hop + 100, 300) - 1
            assert s2[i * n2 + j] == min(r1 // 32 + 2, 7) * 10 + min(c1 // 32 + 2, 9)


def test_bf16_rounding(tn):
    assert tn.bf16(1.0) == 1.0
    assert tn.bf16(1.0 + 2 ** -9) == 1.0           # tie to even
    assert tn.bf16(1.0 + 3 * 2 ** -9) == 1.0 + 2 ** -7
    assert tn.bf16(-3.0e-3) == pytest.approx(-3.0e-3, rel=2 ** -8)


def test_network_input_is_double_centred_and_scaled(tn):
    rng = np.random.default_rng(9)
    P = rng.random((17, 23)) + np.arange(17)[:, None] * 0.1 + np.arange(23)[None, :] * 0.05
    X = tn.network_input(P)
    assert np.allclose(X.mean(axis=0), 0, atol=1e-12)
    assert np.allclose(X.mean(axis=1), 0, atol=1e-12)
    assert np.mean(X * X) == pytest.approx(1.0, rel=1e-9)
    # additive row and column offsets do not change it
    assert np.allclose(tn.network_input(P + rng.random(17)[:, None] + rng.random(23)[None, :]), X)
