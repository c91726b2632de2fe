"""Batches of collective systems: every replica of a batch must end
bit-identical to the same system advanced alone (same Philox key), with the
interval launches replayed from a HIP graph or issued directly, through
re-sorts, with uneven boxes and parameters."""
from __future__ import annotations

import numpy as np
import pytest

from main_amd.batch import CollectiveBatch
from main_amd.collective import CollectiveEngine
from main_amd.tour import cell_tour
from tests.parity_cases import build_collective, small_collective_params

pytestmark = pytest.mark.gpu

REPLICAS = [dict(Lx=16.0, Ly=16.0, Lz=20.0, EbN=0.1, RPab=1.0),
            dict(Lx=20.0, Ly=18.0, Lz=24.0, EbN=0.5, RPab=0.6),
            dict(Lx=14.0, Ly=22.0, Lz=16.0, EbN=1.0, RPab=0.3),
            dict(Lx=24.0, Ly=24.0, Lz=12.0, EbN=0.2, RPab=0.9)]


def _engine(i, kw):
    p = small_collective_params(sort_step=20, k1=0.05, **kw)
    s = build_collective(p, seed=40 + i, flags_mix=(i % 2 == 1))
    e = CollectiveEngine(p, rng_mode=1, seed=1000 + i, overlap=False)   # batches replay the events bitwise
    L = [int(v) for v in p["box"]]
    e.set_state(s, tour=cell_tour(*L))
    e.setup(do_sort=True)
    return e


def _states(engines):
    return [e.get_state() for e in engines]


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_replicas_equal_single_runs(use_graph):
    ncoll = 5   # re-sorts at the 2nd and 4th collision
    alone = []
    for i, kw in enumerate(REPLICAS):
        e = _engine(i, kw)
        e.run(ncoll)
        alone.append(e.get_state())
        e.close()
    engines = [_engine(i, kw) for i, kw in enumerate(REPLICAS)]
    batch = CollectiveBatch(engines, tile=512)
    try:
        batch.run(ncoll, use_graph=use_graph)
        got = _states(engines)
    finally:
        batch.close()
        for e in engines:
            e.close()
    for i, (g, a) in enumerate(zip(got, alone)):
        assert g["cstep"] == a["cstep"] == ncoll * 10
        for k in ("x", "y", "z", "vx", "vy", "vz", "flag", "xd", "vxd", "fxd"):
            assert np.array_equal(g[k], a[k]), f"replica {i}: {k}"


def test_autotune_keeps_results_and_picks_a_tile():
    engines = [_engine(i, kw) for i, kw in enumerate(REPLICAS[:2])]
    ref = [_engine(i, kw) for i, kw in enumerate(REPLICAS[:2])]
    batch = CollectiveBatch(engines, tile=256)
    try:
        times = batch.autotune(tiles=(256, 1024, 4096), ncoll=1)
        assert set(times) == {256, 1024, 4096} and batch.tile in times
        n = 3 * 2   # intervals the autotuner ran
        for e in ref:
            e.run(n)
        for e, r in zip(engines, ref):
            g, a = e.get_state(), r.get_state()
            assert np.array_equal(g["x"], a["x"]) and np.array_equal(g["vz"], a["vz"])
    finally:
        batch.close()
        for e in engines + ref:
            e.close()


def test_body_average_matches_single_engine_diagnostics():
    """Per-replica sums of the batch's body-frame B counts and propulsion
    velocity over three intervals equal those of each system advanced alone
    and sampled after every interval with the single-engine diagnostics."""
    n, grid = 3, dict(nr=16, nz=32, rmax=8.0, zmin=-10.0, zmax=6.0)
    ref_counts, ref_v = [], []
    for i, kw in enumerate(REPLICAS[:3]):
        e = _engine(i, kw)
        c, v = np.zeros((grid["nz"], grid["nr"]), np.int64), 0.0
        for _ in range(n):
            e.run(1)
            c += e.body_density(**grid)[1].astype(np.int64)
            v += e.observables()["v_prop"]
        ref_counts.append(c)
        ref_v.append(v)
        e.close()
    engines = [_engine(i, kw) for i, kw in enumerate(REPLICAS[:3])]
    batch = CollectiveBatch(engines, tile=512)
    try:
        counts, vmean = batch.body_average(n, **grid)
    finally:
        batch.close()
        for e in engines:
            e.close()
    for i in range(3):
        assert np.array_equal(np.rint(counts[i] * n).astype(np.int64), ref_counts[i]), f"replica {i}"
        assert abs(vmean[i] * n - ref_v[i]) <= 1e-12 * (abs(ref_v[i]) + 1e-3)
    assert sum(int(c.sum()) for c in ref_counts) > 200     # replica 1 starts with half its solvent B
