"""Surrogate-guided sweep on the GPU at a small size: batched MPC runs with
per-replica parameters give body-frame B fields that the surrogate, trained on
the GPU, predicts for held-out parameter points better than the method
without it (the training mean), and selection returns the candidates it ranks
highest."""
from __future__ import annotations

import numpy as np
import pytest

from main_amd import sweep as W

pytestmark = pytest.mark.gpu


def test_surrogate_beats_mean_field_on_held_out_runs():
    grid = W.Grid(nr=16, nz=32, rmax=8.0, zmin=-14.0, zmax=8.0)
    rng = np.random.default_rng(5)
    X = W.sample_params(120, rng)
    F, V = W.mpc_runs(X, L=20, warm=40, avg=160, seeds=np.arange(120) + 90, grid=grid, batch_size=120)
    assert np.all(np.isfinite(F)) and np.all(np.isfinite(V))
    assert F.max() > 0.05                     # B is made at the catalytic sphere
    ntr = 96
    sur = W.SweepSurrogate(grid, n_hidden=128, members=2)
    try:
        losses = sur.fit(X[:ntr], F[:ntr], V[:ntr], stages=((300, 1e-3), (200, 3e-4)))
        assert losses[-1] < 0.5 * losses[0]
        assert all(0 < m.best_epoch <= 500 for m in sur.members)
        Fp, Vp = sur.predict(X[ntr:])
        err_sur = W.field_error(Fp, F[ntr:]).mean()
        err_mean = W.field_error(np.broadcast_to(F[:ntr].mean(axis=0), F[ntr:].shape), F[ntr:]).mean()
        print(f"held-out field error: surrogate {err_sur:.4f}, training mean {err_mean:.4f}")
        assert err_sur < 0.7 * err_mean
        C, vp, sd = W.select(sur, 256, 8, np.random.default_rng(1))
        assert C.shape == (8, 8) and np.all(np.diff(vp) <= 0) and np.all(sd >= 0)
        _, vall = sur.predict(C)
        assert np.allclose(vall, vp, rtol=1e-6, atol=1e-9)
        members = sur.predict_members(C)
        assert members.shape == (2, 8, grid.nz * grid.nr + 1)
        assert np.allclose(members.mean(axis=0)[:, -1], vp, rtol=1e-6, atol=1e-9)
    finally:
        sur.close()
