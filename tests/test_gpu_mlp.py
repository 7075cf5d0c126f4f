"""Learned-scheduler cost model on the GPU: matrix-core operand layout,
backward pass against the FP64 oracle (oracle/mlp_ref.py), and training."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("M,N,K", [(16, 16, 16), (37, 21, 50), (128, 64, 300), (5, 64, 4096)])
def test_mfma_gemm_layout(xgm, M, N, K):
    from oracle.mlp_ref import bf16
    _say("mfma gemm", M, N, K)
    rng = np.random.default_rng(M * 1000 + N)
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    got = xgm.mfma_gemm(A, B)
    want = bf16(A) @ bf16(B)
    # BF16 products are exact in FP32; only FP32 accumulation order differs
    assert np.max(np.abs(got - want)) <= 1e-4 * np.sqrt(K) * np.max(np.abs(want)) + 1e-4


def test_mlp_grads_match_fp64_oracle(xgm):
    from oracle import mlp_ref
    _say("mlp grads")
    rng = np.random.default_rng(3)
    n = 777
    X = rng.standard_normal((n, 16)).astype(np.float32)
    T = rng.standard_normal((n, 5)).astype(np.float32)
    Mk = (rng.random((n, 5)) < 0.7).astype(np.float32)
    m = xgm.MLP(0, seed=11)
    p = m.params
    loss_gpu, g_gpu = m.grads(X, T, Mk)
    loss_ref, g_ref = mlp_ref.loss_and_grads(p.astype(np.float64), X, T, Mk)
    assert abs(loss_gpu - loss_ref) <= 1e-3 * abs(loss_ref)
    # per tensor: FP32 accumulation vs FP64, same BF16 rounding points
    o = 0
    for name, shape in mlp_ref.SIZES:
        k = int(np.prod(shape))
        a, b = g_gpu[o:o + k], g_ref[o:o + k]
        assert np.max(np.abs(a - b)) <= 2e-3 * np.max(np.abs(b)) + 1e-6, name
        o += k
    # a planted sign error in the oracle must be caught by this tolerance
    assert np.max(np.abs(g_gpu[:1024] + g_ref[:1024])) > 2e-3 * np.max(np.abs(g_ref[:1024]))


def test_mlp_trains_and_predicts(xgm):
    _say("mlp train")
    rng = np.random.default_rng(5)
    n = 20000
    X = rng.standard_normal((n, 16)).astype(np.float32)
    W = rng.standard_normal((16, 5)).astype(np.float32) * 0.5
    T = np.tanh(X @ W) + 0.1 * X[:, :5] ** 2
    Mk = np.ones((n, 5), np.float32)
    m = xgm.MLP(0, seed=2)
    y0 = m.predict(X)
    l0 = float(np.mean((y0 - T) ** 2))
    l1 = m.train(X, T, Mk, epochs=30, batch=1024, lr=3e-3)
    y1 = m.predict(X)
    mse = float(np.mean((y1 - T) ** 2))
    assert mse < 0.1 * l0, (l0, l1, mse)
    # set/get round trip
    p = m.params
    m2 = xgm.MLP(0, seed=99)
    m2.params = p
    assert np.allclose(m2.predict(X[:100]), m.predict(X[:100]))
