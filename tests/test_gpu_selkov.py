"""GPU parity and properties of the Selkov engine (selkov_code).

Parity (host mt95 draws in the reference's call order): the engine and the
oracle (pinned bit-exact to the reference's Fortran modules) advance the same
seeded system through whole collision intervals: flags and particle order
must be identical; positions, velocities, forces and the dimer within 1e-12
relative (scales: 1 for positions, sqrt(kT/m) for velocities, 1 for forces;
the dimer force is a tree sum where the reference sums in particle order).
"""
from __future__ import annotations

import copy
import sys

import numpy as np
import pytest

from main_amd import systems as S
from main_amd._lib import Mt95Stream
from main_amd.selkov import SelkovEngine
from oracle import oracle as O
from tests.parity_cases import close

pytestmark = pytest.mark.gpu
SEED = 8154543


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def small_selkov(**kw):
    base = dict(L=20.0, rho=4.0, Rc=2.0, Rn=4.0, Dcn=7.0, MDt=0.005, MPCt=0.1, NG=3200,
                k=(0.05, 0.01, 0.0004, 0.0004, 0.01, 0.01))
    base.update(kw)
    return S.selkov_params(**base)


def selkov_parity_case(ncoll=2, seed=3, p=None, sort_step=10000, flags_mix=True):
    p = p or small_selkov()
    s = S.selkov_initial_state(p, seed=seed)
    if flags_mix:
        s["flag"][:] = np.random.default_rng(seed + 6).choice([0, 1, 3], p["NT"]).astype(np.int64)
    ref = copy.deepcopy(s)
    ref["sort_step"] = sort_step
    O.sk_run(ref, 1, ncoll * p["MPCtstep"], seed=SEED)
    eng = SelkovEngine(p, rng_mode=0, sort_step=sort_step)
    try:
        eng.set_state(s, cstep=0)
        stream = Mt95Stream(SEED)
        eng.run(ncoll, stream)
        g = eng.get_state()
    finally:
        eng.close()
    res = {}
    vs = np.sqrt(p["T"] / p["Mf"])
    res["flag"] = (int(np.sum(g["flag"] != ref["flag"])), 0.0)
    for k in ("x", "y", "z"):
        res[k] = close(g[k], ref[k], 1.0)
    for k in ("vx", "vy", "vz"):
        res[k] = close(g[k], ref[k], vs)
    for k in ("fx", "fy", "fz"):
        res[k] = close(g[k], ref[k], 1.0)
    dm_ref = np.asarray(ref["dm"])
    res["dimer_x"] = close(g["dm"][:, 0:3], dm_ref[:, 0:3], 1.0)
    res["dimer_v"] = close(g["dm"][:, 3:6], dm_ref[:, 3:6], 1e-3)
    res["dimer_f"] = close(g["dm"][:, 6:9], dm_ref[:, 6:9], 1.0)
    res["NT"] = (int(g["NT"] != ref["NT"]), 0.0)
    bad = {k: v for k, v in res.items() if v[0] != 0}
    worst = max(v[1] for v in res.values())
    summary = f"NT={p['NT']} ncoll={ncoll} worst_rel={worst:.2e} draws={g['used']} mismatches={bad or 'none'}"
    return dict(ok=not bad, summary=summary, gpu=g, ref=ref)


@pytest.mark.parametrize("ncoll,seed", [(1, 3), (3, 4)])
def test_selkov_engine_matches_oracle(ncoll, seed):
    _say(f"[selkov] ncoll={ncoll} seed={seed}")
    r = selkov_parity_case(ncoll=ncoll, seed=seed)
    _say(f"[selkov] {r['summary']}")
    assert r["ok"], r["summary"]


def test_selkov_engine_matches_oracle_through_a_sort():
    """Sort.f90 (cell-ordered, species-grouped, compacting) at the second collision."""
    p = small_selkov()
    _say("[selkov] sort at step 40")
    r = selkov_parity_case(ncoll=3, seed=5, p=p, sort_step=2 * p["MPCtstep"])
    _say(f"[selkov] sort: {r['summary']}")
    assert r["ok"], r["summary"]


def test_selkov_engine_matches_oracle_dense_fast_reactions():
    p = small_selkov(rho=8.0, NG=6000, k=(0.5, 0.1, 0.004, 0.004, 0.1, 0.1))
    r = selkov_parity_case(ncoll=2, seed=7, p=p)
    _say(f"[selkov] dense: {r['summary']}")
    assert r["ok"], r["summary"]


def test_selkov_throughput_mode_properties():
    """50^3 cells at 10 per cell (1,250,000 particles), device Philox draws,
    a sort every third interval: particle number, flags, confinement,
    equipartition, bit-identical reruns, and bit-identical states with the far
    particles streamed beside the near engine (overlap) and after it."""
    p = S.selkov_params(L=50.0, rho=10.0, NG=100000)
    outs = []
    for overlap in (True, True, False):
        _say(f"[selkov] throughput mode, overlap={overlap}")
        s = S.selkov_initial_state(p, seed=17)
        eng = SelkovEngine(p, rng_mode=1, seed=99, sort_step=3 * p["MPCtstep"])
        try:
            eng.set_overlap(overlap)
            eng.set_state(s, cstep=0)
            eng.run(6)
            outs.append(eng.get_state())
        finally:
            eng.close()
    g = outs[0]
    assert g["cstep"] == 6 * p["MPCtstep"] and g["NT"] == p["NT"]
    assert set(np.unique(g["flag"])) <= {0, 1, 3}
    L = p["box"][0]
    for k in ("x", "y", "z"):
        assert np.all((g[k] >= 0) & (g[k] < L))
    T = (p["Mf"] * (g["vx"] ** 2 + g["vy"] ** 2 + g["vz"] ** 2)).sum() / (3.0 * p["NT"])
    assert abs(T / p["T"] - 1.0) < 5e-3, T
    for o in outs[1:]:
        for k in ("x", "y", "z", "vx", "vy", "vz", "flag", "fx", "dm"):
            assert np.array_equal(outs[0][k], o[k]), k
