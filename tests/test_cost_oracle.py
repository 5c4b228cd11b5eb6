"""The cost-build restatements (oracle C code and the NumPy fixed-order form
used for golden inputs), pinned by closed forms and by an independent
summation order.  C = 1 - L2norm(F1)^T L2norm(F2) with L2norm(F) =
F / (sqrt(sum_k F^2) + 1e-9), the epsilon added after the square root
(FlexDTW.py:1207-1210, Parflex.py:1470).  The reference's own product is a
BLAS dgemm whose summation order is unspecified, so bit equality with it is
not defined; agreement with NumPy's matmul is checked to 1e-14."""
import numpy as np
import pytest


def test_two_restatements_agree_bitwise(oracle_mod):
    from parflexdtw_amd.synth import cost_matrix_det, make_pair
    for K, N, M, seed in [(12, 70, 90, 1), (128, 33, 40, 2), (3, 5, 200, 3)]:
        F1, F2, _ = make_pair(N, M + 10, K=K, seed=seed, kind="relu")
        a = cost_matrix_det(F1, F2)
        b = oracle_mod.cost_matrix(F1, F2)
        assert np.array_equal(a, b)
        ref = 1.0 - (F1 / (np.linalg.norm(F1, axis=0) + 1e-9)).T @ (
            F2 / (np.linalg.norm(F2, axis=0) + 1e-9))
        assert np.max(np.abs(a - ref)) < 1e-14


def test_closed_forms(oracle_mod):
    from parflexdtw_amd.synth import cost_matrix_det
    K = 6
    # one-hot frames scaled by s: C = 1 - [k_i == k_j] * s^2 / (s + 1e-9)^2
    ki = np.array([0, 1, 2, 3, 4, 5, 0, 3])
    kj = np.array([5, 0, 3, 3, 1])
    si = np.array([1.0, 2.0, 0.5, 1e-8, 3.0, 1.0, 1e-9, 7.0])
    sj = np.array([1.0, 4.0, 1e-8, 2.0, 1.0])
    F1 = np.zeros((K, ki.size))
    F1[ki, np.arange(ki.size)] = si
    F2 = np.zeros((K, kj.size))
    F2[kj, np.arange(kj.size)] = sj
    want = 1.0 - (ki[:, None] == kj[None, :]) * (si / (si + 1e-9))[:, None] * \
        (sj / (sj + 1e-9))[None, :]
    for C in (cost_matrix_det(F1, F2), oracle_mod.cost_matrix(F1, F2)):
        assert np.allclose(C, want, rtol=0, atol=1e-15)
    # epsilon after the square root: a frame of norm 1e-8 keeps ~0.909 of
    # its weight (inside the root it would keep ~0.3)
    assert C[3, 2] == pytest.approx(1.0 - (1e-8 / (1e-8 + 1e-9)) ** 2, abs=1e-15)
    # an all-zero frame has cost exactly 1
    F1[:, 0] = 0.0
    assert np.all(oracle_mod.cost_matrix(F1, F2)[0] == 1.0)
    assert np.all(cost_matrix_det(F1, F2)[0] == 1.0)
