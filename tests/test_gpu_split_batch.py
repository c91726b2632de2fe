"""Split-precision batches of collective replicas (mpcx_cob_* with split
engines).  Every replica of a batch ends bit-identical to the same system run
alone (a split engine's own run is a batch of one), with the interval
replayed from a HIP graph or issued directly, through re-sorts, with uneven
boxes and per-replica parameters.  The ensemble workload of 256 replicas with
box edges 20-40 cells runs at full size; a subset of it against FP64 runs of
the same systems compares the propulsion velocities."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from main_amd import systems as S
from main_amd.batch import CollectiveBatch
from main_amd.collective import CollectiveEngine
from main_amd.tour import cell_tour

pytestmark = pytest.mark.gpu

_TOURS: dict = {}


def _tour(L):
    if L not in _TOURS:
        _TOURS[L] = cell_tour(L, L, L)
    return _TOURS[L]


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _params(L, EbN=0.1, RPab=1.0, sort_step=500):
    return S.collective_params(Lx=float(L), Ly=float(L), Lz=float(L), numdims=1, rhoa=10.0, Rc=2.0, Rn=4.0,
                               Dcn=6.5, EbN=EbN, RPab=RPab, T=1.0 / 6.0, skinrad=0.3, MDt=0.01, MPCt=0.5,
                               sort_step=sort_step, k1=1e-3)


def _engine(p, seed, split=True, host_state=True):
    L = int(p["box"][0])
    e = CollectiveEngine(p, rng_mode=1, seed=seed, split=split)
    if host_state:
        e.set_state(S.collective_initial_state(p, seed=seed), tour=_tour(L))
    else:
        e.set_dimers(S.collective_dimers(p, seed=seed), tour=_tour(L))
        e.init_solvent(seed)
    e.setup(do_sort=True)
    return e


REPLICAS = [(20, 0.1, 1.0), (26, 0.5, 0.6), (23, 1.0, 0.3), (30, 0.2, 0.9)]
KEYS = ("x", "y", "z", "vx", "vy", "vz", "flag", "xd", "yd", "zd", "vxd", "vyd", "vzd", "fxd")


# (graph, tile, binning table slots): 1024 slots by default below 4096-particle
# tiles, 2048 from there; the single runs use 2048-particle tiles (1024 slots)
@pytest.mark.parametrize("use_graph,tile,slots", [(True, 512, 0), (False, 512, 0), (True, 4096, 0),
                                                  (False, 1024, 2048)])
def test_split_batch_replicas_equal_single_runs(use_graph, tile, slots):
    ncoll = 5   # re-sorts at the 2nd and 4th collision
    ps = [_params(L, EbN, RPab, sort_step=100) for L, EbN, RPab in REPLICAS]
    alone = []
    for i, p in enumerate(ps):
        e = _engine(p, 70 + i)
        e.run(ncoll)
        alone.append(e.get_state())
        e.close()
    engines = [_engine(p, 70 + i) for i, p in enumerate(ps)]
    batch = CollectiveBatch(engines, tile=tile)
    try:
        if slots:
            batch.set_bulk_slots(slots)
        batch.run(ncoll, use_graph=use_graph)
        got = [e.get_state() for e in engines]
    finally:
        batch.close()
        for e in engines:
            e.close()
    for i, (g, a) in enumerate(zip(got, alone)):
        assert g["cstep"] == a["cstep"] == ncoll * 50
        for k in KEYS:
            assert np.array_equal(g[k], a[k]), f"replica {i}: {k}"


def test_split_batch_sort_period_knob_equals_engines_at_that_period():
    """A split batch's re-sort period set through mpcx_cob_set_sort_step gives
    the states of engines configured with that period run alone; FP64
    batches refuse the knob."""
    ncoll = 5   # re-sort at the 3rd collision (150 MD steps)
    alone = []
    for i, (L, EbN, RPab) in enumerate(REPLICAS[:3]):
        e = _engine(_params(L, EbN, RPab, sort_step=150), 90 + i)
        e.run(ncoll)
        alone.append(e.get_state())
        e.close()
    engines = [_engine(_params(L, EbN, RPab, sort_step=100), 90 + i) for i, (L, EbN, RPab) in enumerate(REPLICAS[:3])]
    batch = CollectiveBatch(engines, tile=1024)
    try:
        batch.set_sort_step(150)
        batch.run(ncoll)
        got = [e.get_state() for e in engines]
    finally:
        batch.close()
        for e in engines:
            e.close()
    for i, (g, a) in enumerate(zip(got, alone)):
        for k in KEYS:
            assert np.array_equal(g[k], a[k]), f"replica {i}: {k}"
    fp = [_engine(_params(20, sort_step=100), 95, split=False)]
    b64 = CollectiveBatch(fp, tile=1024)
    try:
        with pytest.raises(Exception, match="sort_step"):
            b64.set_sort_step(150)
    finally:
        b64.close()
        fp[0].close()


def test_split_batch_overflowing_inboxes_equal_single_runs():
    """Inboxes of two entries: nearly every sum another tile binned into a
    cell goes through the atomics into the accumulator copies instead (the
    overflow path), and every replica still ends bitwise equal to its engine
    run alone."""
    ncoll = 4   # a re-sort at the 2nd collision
    ps = [_params(L, EbN, RPab, sort_step=100) for L, EbN, RPab in REPLICAS[:3]]
    alone = []
    for i, p in enumerate(ps):
        e = _engine(p, 110 + i)
        e.run(ncoll)
        alone.append(e.get_state())
        e.close()
    engines = [_engine(p, 110 + i) for i, p in enumerate(ps)]
    batch = CollectiveBatch(engines, tile=1024)
    try:
        batch.set_inbox_cap(2)
        batch.run(ncoll, use_graph=True)
        got = [e.get_state() for e in engines]
    finally:
        batch.close()
        for e in engines:
            e.close()
    for i, (g, a) in enumerate(zip(got, alone)):
        for k in KEYS:
            assert np.array_equal(g[k], a[k]), f"replica {i}: {k}"


def _ensemble(R, seed):
    rng = np.random.default_rng(seed)
    return [(int(rng.integers(20, 41)), float(rng.uniform(0.1, 1.0)), float(rng.uniform(0.2, 1.0)))
            for _ in range(R)]


def test_ensemble_of_256_replicas_runs_and_matches_single_runs():
    """The replica ensemble at full size: 256 systems, box edges uniform on
    20..40 cells, 10 per cell, EbN in [0.1, 1], RPab in [0.2, 1], one key
    each, split storage, graph replay, a re-sort every 10 collisions.  After
    12 intervals every replica is intact (temperature within 2 % of kT,
    particles inside its box) and three of them equal runs alone."""
    cfg = _ensemble(256, 4357787)
    ps = [_params(L, EbN, RPab) for L, EbN, RPab in cfg]
    engines = []
    for r, p in enumerate(ps):
        engines.append(_engine(p, 4357787 + r, host_state=False))
    npart = sum(int(p["NT"]) for p in ps)
    _say(f"[ensemble] 256 replicas, {npart} particles set up")
    batch = CollectiveBatch(engines, tile=2048)
    picks = (0, 117, 255)
    try:
        batch.run(12, use_graph=True)
        _say(f"[ensemble] 12 intervals in {batch.info()['run_ms']:.1f} ms")
        got = {}
        for r, (e, p) in enumerate(zip(engines, ps)):
            o = e.observables()
            assert abs(o["T_sol"] / p["T"] - 1.0) < 0.02, f"replica {r}: T {o['T_sol']}"
            if r in picks:
                got[r] = e.get_state()
                assert np.all((got[r]["z"] >= 0.0) & (got[r]["z"] < p["box"][2]))
    finally:
        batch.close()
        for e in engines:
            e.close()
    for r in picks:
        e = _engine(ps[r], 4357787 + r, host_state=False)
        try:
            e.run(12)
            a = e.get_state()
        finally:
            e.close()
        for k in KEYS:
            assert np.array_equal(got[r][k], a[k]), f"replica {r}: {k}"


def _mean_velocity(engines, batch, n, burn):
    batch.run(burn, use_graph=True)
    R = len(engines)
    v = np.zeros((n, R))
    for t in range(n):
        batch.run(1, use_graph=True)
        v[t] = [e.observables()["v_prop"] for e in engines]
    return v


def test_split_ensemble_propulsion_tracks_fp64():
    """16 replicas of the ensemble (their boxes and parameters) in split and in
    FP64 storage from the same initial states and keys: the ensemble-mean
    propulsion velocity over 800 collisions agrees within 2 sigma (standard
    error of the difference from the replicas' time means; deterministic
    runs: the seeds fix the outcome).  The FP64 engine stands in for the
    oracle here: it follows the reference bit for bit in the integer state and
    to 1e-12 in the floating-point state per interval (test_gpu_collective),
    and the CPU oracle cannot run 16 x 800 collisions of these boxes in a test."""
    cfg = _ensemble(256, 4357787)[:16]
    means = {}
    for split in (False, True):
        ps = [_params(L, EbN, RPab) for L, EbN, RPab in cfg]
        engines = [_engine(p, 4357787 + r, split=split, host_state=False) for r, p in enumerate(ps)]
        batch = CollectiveBatch(engines, tile=2048)
        try:
            v = _mean_velocity(engines, batch, 800, 50)
        finally:
            batch.close()
            for e in engines:
                e.close()
        means[split] = v.mean(axis=0)
        _say(f"[propulsion] split={split}: ensemble mean {v.mean():.5f}")
    d = means[True] - means[False]
    se = d.std(ddof=1) / np.sqrt(d.size)
    _say(f"[propulsion] mean difference {d.mean():.5f}, 2 sigma {2 * se:.5f}")
    assert abs(d.mean()) < 2.0 * se
    assert means[False].mean() > 0.0   # the motors do propel (N sphere attracts B less)
