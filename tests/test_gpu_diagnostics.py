"""GPU diagnostics against the checker (tests/diag_reference.py, pinned to the
reference's own output routines by tests/test_diag_pins.py) on the same
device state: counts bit-exact; sums within 1e-12 of the magnitudes summed
(the GPU adds in a tree, the reference in particle order); the dimer terms
(no particle sums) bit-exact."""
from __future__ import annotations

import numpy as np
import pytest

from main_amd import systems as S
from main_amd.collective import CollectiveEngine
from main_amd.selkov import SelkovEngine
from tests import diag_reference as R
from tests.parity_cases import build_collective, small_collective_params

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _check_obs(o, r):
    assert o["n"] == r["n"]
    assert [o[k] for k in list(o)[1:4]] == r["counts"]
    scale = r["pabs"]
    for i, k in enumerate(("psol_x", "psol_y", "psol_z")):
        assert abs(o[k] - r["psol"][i]) <= TOL * scale
    for i, k in enumerate(("p_x", "p_y", "p_z")):
        assert abs(o[k] - r["ptot"][i]) <= TOL * scale
    for k in ("ke_sol", "T_sol", "E_tot"):
        assert abs(o[k] - r[k]) <= TOL * max(abs(r[k]), 1.0), k
    # no particle sums in these: the reference's expressions, its 2./3. (single precision)
    assert o["ke_dim"] == r["ke_dim"] and o["T_dim"] == r["T_dim"]
    assert abs(o["T_sol"] / o["ke_sol"] * o["n"] - R.TWO_THIRDS) <= 1e-15
    assert o["v_prop"] == r["v_prop"]


def test_collective_diagnostics():
    p = small_collective_params(sort_step=20, k1=0.2)
    s = build_collective(p, seed=31, flags_mix=True)
    eng = CollectiveEngine(p, rng_mode=1, seed=5)
    try:
        eng.set_state(s, tour=np.arange(1, p["vol"] + 1))
        eng.setup(do_sort=False)
        eng.run(3)
        g = eng.get_state()
        o = eng.observables()
        sl = eng.conc_slabs(8.0, 8.0, 3.0, p["Rc"] * p["cutoff"])
        bd = eng.body_density(nr=16, nz=32, rmax=8.0, zmin=-10.0, zmax=6.0)
        tp = eng.get_cells()["tpotential"]
    finally:
        eng.close()
    nd = p["numdims"]
    st = dict(g)
    st["xd"] = np.stack([np.r_[g["xd"][:, 0], g["xd"][:, 1]], np.r_[g["yd"][:, 0], g["yd"][:, 1]],
                         np.r_[g["zd"][:, 0], g["zd"][:, 1]]], axis=1)
    st["vd"] = np.stack([np.r_[g["vxd"][:, 0], g["vxd"][:, 1]], np.r_[g["vyd"][:, 0], g["vyd"][:, 1]],
                         np.r_[g["vzd"][:, 0], g["vzd"][:, 1]]], axis=1)
    r = R.observables(st, {0: p["Ma"], 1: p["Mb"], 2: p["Ms"]}, {0: 0, 1: 1, 2: 2}, (p["Mc"], p["Mn"]), nd,
                      p["box"], False, tp)
    _check_obs(o, r)
    rs = R.slabs(g, p["box"], 8.0, 8.0, 3.0, p["Rc"] * p["cutoff"], 0.5)
    for k in ("A1d", "B1d", "Axy", "Bxy", "Axz", "Bxz", "Ayz", "Byz"):
        assert np.array_equal(sl[k], rs[k]), k
    assert sl["B1d"].sum() > 0
    rb = R.body_density(g, {0: 0, 1: 1, 2: 2}, p["box"], False, st["xd"][0], r["axis"], 16, 32, 8.0, -10.0, 6.0)
    assert np.array_equal(bd.astype(np.int64), rb)
    assert bd.sum() > 1000


def test_selkov_diagnostics():
    p = S.selkov_params(L=20.0, rho=4.0, NG=3200, k=(0.05, 0.01, 0.0004, 0.0004, 0.01, 0.01))
    s = S.selkov_initial_state(p, seed=8)
    eng = SelkovEngine(p, rng_mode=1, seed=3)
    try:
        eng.set_state(s, cstep=0)
        eng.run(3)
        g = eng.get_state()
        o = eng.observables()
        sh = eng.shells()
        bd = eng.body_density(nr=16, nz=32, rmax=8.0, zmin=-12.0, zmax=6.0)
    finally:
        eng.close()
    dm = g["dm"]
    st = dict(g, xd=dm[:, 0:3].copy(), vd=dm[:, 3:6].copy())
    r = R.selkov_observables(st, {0: p["Mf"], 1: p["Mg"], 3: p["Mi"]}, (p["Mc"], p["Mn"]), p["box"],
                             g["tpotential"])
    _check_obs(o, r)
    assert np.array_equal(sh, R.shells(g, p["box"], dm[0, 0:3]))
    assert sh.sum() > 100
    rb = R.body_density(g, {0: 0, 1: 1, 3: 2}, p["box"], True, dm[0, 0:3], r["axis"], 16, 32, 8.0, -12.0, 6.0)
    assert np.array_equal(bd.astype(np.int64), rb)
