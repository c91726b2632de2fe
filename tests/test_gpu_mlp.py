"""The surrogate's BF16 MFMA kernels against the numpy oracle
(tests/mlp_reference.py, pinned by finite differences): the GEMM, the loss
and every parameter gradient of one batch, and a training run whose loss
must fall."""
import numpy as np
import pytest

from main_amd.surrogate import Surrogate, gemm_bf16
from tests import mlp_reference as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(64, 64, 32), (100, 70, 45), (256, 2049, 256), (2049, 256, 96)])
def test_gemm_bf16_matches_reference(shape):
    m, n, k = shape
    rng = np.random.default_rng(m + n + k)
    A = rng.normal(size=(m, k)).astype(np.float32)
    B = rng.normal(size=(k, n)).astype(np.float32)
    Cg, _ = gemm_bf16(A, B)
    Cr = M.to_bf16(A) @ M.to_bf16(B)
    scale = np.abs(M.to_bf16(A)) @ np.abs(M.to_bf16(B))
    assert np.max(np.abs(Cg - Cr) / scale) < 1e-6   # FP32 accumulation of exact BF16 products


def test_loss_and_gradients_match_reference():
    rng = np.random.default_rng(1)
    net = Surrogate(n_in=8, n_hidden=64, n_layers=3, n_out=33, seed=7)
    try:
        flat = net.params().astype(np.float64)
        x = rng.normal(size=(50, 8)).astype(np.float32)
        t = rng.normal(size=(50, 33)).astype(np.float32)
        w = rng.uniform(0.5, 1.5, size=33).astype(np.float32)
        loss, g = net.loss_grad_raw(x, t, w)
        lr, gr = M.loss_grad(flat, net.dims(), x.astype(np.float64), t.astype(np.float64), w.astype(np.float64),
                             bf16=True)
        assert abs(loss - lr) < 1e-5 * lr
        assert np.linalg.norm(g - gr) < 2e-3 * np.linalg.norm(gr)
        # every layer's block, separately
        o = 0
        d = net.dims()
        for a, b in zip(d[:-1], d[1:]):
            for sz in (a * b, b):
                blk = slice(o, o + sz)
                assert np.linalg.norm(g[blk] - gr[blk]) < 5e-3 * np.linalg.norm(gr[blk]) + 1e-7
                o += sz
        y = net.forward_raw(x)
        H, _ = M.forward(flat, d, x.astype(np.float64), bf16=True)
        assert np.max(np.abs(y - H[-1])) < 2e-3 * np.max(np.abs(H[-1]))
    finally:
        net.close()


def test_training_reduces_loss():
    rng = np.random.default_rng(2)
    X = rng.uniform(-1, 1, size=(512, 8))
    P = rng.normal(size=(8, 65))
    Y = np.sin(X @ P) + 0.1 * (X ** 2).sum(axis=1, keepdims=True)
    net = Surrogate(n_in=8, n_hidden=256, n_layers=3, n_out=65, seed=3)
    try:
        losses = net.train_raw(X, Y, np.full(65, 1.0 / 65), batch=64, epochs=60, lr=1e-3, seed=1)
        assert losses[-1] < 0.1 * losses[0], losses[[0, -1]]
        assert np.all(np.isfinite(losses))
    finally:
        net.close()
