"""Selkov text records and restart files written from the device engine.

The state the engine holds after two collision intervals is written in the
reference's FORMAT layouts (Output_SelkovDimer.f90), read back the way
Restart.f90 reads it (species order F, G, I) and uploaded into a new engine:
the solvent comes back bit for bit (ES23.16 has 17 significant digits), the
dimer to the 16 digits of E23.16.  The DotVZ and energy/momentum records
agree with the engine's own device observables.
"""
from __future__ import annotations

import numpy as np
import pytest

from main_amd import selkov_files as SF
from main_amd import systems as S
from main_amd.selkov import SelkovEngine
from tests.test_gpu_selkov import small_selkov

pytestmark = pytest.mark.gpu


def _engine(p, seed=3):
    s = S.selkov_initial_state(p, seed=seed)
    s["flag"][:] = np.random.default_rng(seed + 6).choice([0, 1, 3], p["NT"]).astype(np.int64)
    eng = SelkovEngine(p, rng_mode=1, seed=seed)
    eng.set_state(s, cstep=0)
    eng.run(2)
    return eng


def test_selkov_restart_files_round_trip(tmp_path):
    p = small_selkov()
    eng = _engine(p)
    try:
        cstep = SF.save_restart(eng, str(tmp_path))
        g = eng.get_state()
    finally:
        eng.close()
    assert cstep == 2 * p["MPCtstep"]
    r = SF.read_restart(str(tmp_path), p)
    NT = int(g["NT"])
    flag = g["flag"][:NT]
    order = np.concatenate([np.nonzero(flag == f)[0] for f in (0, 1, 3)])
    assert r["NT"] == NT and r["cstep"] == cstep
    assert np.array_equal(r["flag"], flag[order])
    for k in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz"):
        assert np.array_equal(r[k].view(np.uint64), g[k][:NT][order].view(np.uint64)), k
    want = np.vectorize(lambda v: float("%.15e" % v))(g["dm"][:, :9])
    assert np.array_equal(r["dm"][:, :9], want)
    # the restored state uploads and runs
    e2 = SelkovEngine(p, rng_mode=1, seed=4)
    try:
        e2.set_state(r, cstep=cstep)
        g2 = e2.get_state()
        for k in ("x", "vx", "fz"):
            assert np.array_equal(g2[k][:NT], r[k]), k
        e2.run(1)
        assert e2.scalars()["cstep"] == cstep + p["MPCtstep"]
    finally:
        e2.close()


def test_selkov_records_agree_with_device_observables(tmp_path):
    p = small_selkov()
    eng = _engine(p, seed=5)
    try:
        files = SF.output_records(eng, p)
        o = eng.observables()
        g = eng.get_state()
    finally:
        eng.close()
    SF.write_files(str(tmp_path), files)
    cstep = int(g["cstep"])
    dv = SF._numbers(files["DotVZ.dat"])
    assert dv[0] == cstep
    assert abs(dv[1] - o["v_prop"]) <= 1e-14 * max(abs(o["v_prop"]), 1e-3)
    em = SF._numbers(files["EnergyMomentum.dat"])
    assert em[0] == cstep
    for i, k in enumerate(("p_x", "p_y", "p_z")):          # particle-order sums vs device tree sums
        assert abs(em[1 + i] - o[k]) <= 1e-12 * max(abs(o[k]), 1.0), k
    # the device observables follow Output_SelkovDimer.f90:243-300 as the record does: the
    # dimer as its centre of mass, temperatures with the single-precision 2./3.
    assert abs(em[4] - (o["ke_sol"] + o["ke_dim"])) <= 1e-12 * em[4]
    assert abs(em[7] - o["T_sol"]) <= 1e-12 * o["T_sol"]
    assert em[8] == o["T_dim"]
    conc = SF._numbers(files["Concentration.dat"])
    assert conc[1] + conc[2] + conc[3] == int(g["NT"])
    assert 0 < conc[10] <= p["vol"]
    names = {SF.dump_name(f"Solvent{sp}{k}", cstep) for sp in "FGI" for k in ("Pos", "Vel", "For")}
    assert names <= set(files)
