"""Pins the numpy surrogate oracle: its backward pass against central finite
differences of its own FP64 loss, and its BF16 rounding against the IEEE
definition."""
import numpy as np

from tests import mlp_reference as M


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(0)
    dims = [5, 7, 6, 4]
    n = sum(a * b + b for a, b in zip(dims[:-1], dims[1:]))
    flat = rng.normal(size=n) * 0.5
    x = rng.normal(size=(9, 5))
    t = rng.normal(size=(9, 4))
    w = rng.uniform(0.5, 2.0, size=4)
    _, g = M.loss_grad(flat, dims, x, t, w)
    h = 1e-6
    fd = np.empty(n)
    for i in range(n):
        e = np.zeros(n)
        e[i] = h
        fd[i] = (M.loss_grad(flat + e, dims, x, t, w)[0] - M.loss_grad(flat - e, dims, x, t, w)[0]) / (2 * h)
    assert np.max(np.abs(fd - g)) < 1e-6 * max(1.0, np.max(np.abs(g)))


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, -3.14159, 0.0], np.float32)
    r = M.to_bf16(v)
    # 1 + 2^-9 is halfway between 1 and 1 + 2^-7... bf16 keeps 7 fraction bits: spacing 2^-7 near 1
    assert r[0] == 1.0 and r[3] == 1.0                       # below half: down; exactly half: to even (1.0)
    assert r[2] == 1.0 + 2 ** -7                              # 1.5 ulp: up
    assert abs(r[4] - (-3.140625)) < 1e-12
    assert r[5] == 0.0
