"""Pins of the diagnostics checker (tests/diag_reference.py) and of the Selkov
record functions (main_amd/selkov_files.py) to the reference's own output
routines, compiled by oracle/Makefile into oracle/_ref:

* Output_SelkovDimer.f90:147-305 (DotVZ, Concentration, EnergyMomentum) and
  Output_Concentration_Catalytic.f90:7-191 (shells) run on a random state in a
  temporary directory; their ES23.16 records carry 17 significant digits, so
  the values read back are the reference's doubles and are compared exactly.
* the solvent part of the collective observables (output_mod.f95:696-738,
  which needs HDF5 and is not built): the same expressions as the Selkov
  routine's solvent loops, so it is checked against that routine on a state
  relabelled A, B, S -> F, G, I with the dimer at rest (then sumPX is the
  solvent sum and Ekin the solvent kinetic energy).
* the collective dimer terms (output_mod.f95:719-732): against exact rational
  arithmetic, within a few ulps.

The reference-run tests skip when oracle/_ref is not built (it needs the
reference sources, present only where the build ran).
"""
from __future__ import annotations

import ctypes as C
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O
from tests import diag_reference as R

VOL = 64000   # Global_SelkovDimer.f90:30 fixes the reference's Vol (a 40^3 box)


def _state(seed, n=30000, flags=(0, 1, 3, 2), box=40.0, dimer_across=False, dimer_rest=False):
    rng = np.random.default_rng(seed)
    st = dict(x=rng.uniform(0, box, n), y=rng.uniform(0, box, n), z=rng.uniform(0, box, n),
              vx=rng.normal(0, 0.45, n), vy=rng.normal(0, 0.45, n), vz=rng.normal(0, 0.45, n),
              flag=rng.choice(np.array(flags, np.int64), n, p=[0.45, 0.35, 0.15, 0.05][:len(flags)]
                              if len(flags) == 4 else None))
    dm = np.zeros((2, 12), order="F")
    c = np.array([box / 2, box / 2, box / 2]) if not dimer_across else np.array([1.5, box - 2.0, 3.0])
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    nn = (c - 6.5 * u) % box
    dm[0, 0:3], dm[1, 0:3] = c, nn
    if not dimer_rest:
        dm[:, 3:6] = rng.normal(0, 0.05, (2, 3))
    st["dm"] = dm
    st["xd"], st["vd"] = dm[:, 0:3].copy(), dm[:, 3:6].copy()
    return st


def _ref_output(tmp_path, st, masses, msph, tpot, cells, cstep=10):
    lib = O.ref_lib("selkov")
    if lib is None:
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    n = st["x"].size
    ip = np.array([n, n], np.int64)
    rp = np.array([40.0, 40.0, 40.0, masses[0], masses[1], masses[3], msph[0], msph[1], 6.5, tpot])
    freq = np.array([cstep, 7, 7, cstep, cstep, cstep, 7], np.int64)
    arrs = [np.ascontiguousarray(st[k], dtype=np.float64) for k in ("x", "y", "z", "vx", "vy", "vz")]
    flg = np.ascontiguousarray(st["flag"], dtype=np.int64)
    dm = np.asfortranarray(st["dm"])
    cl = [np.ascontiguousarray(cells[k], dtype=np.int64) for k in ("tF", "tG", "tI", "aff")]
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        lib.ref_sk_output(O._ptr(ip), O._ptr(rp), O._ptr(freq), O._ptr(flg), *[O._ptr(a) for a in arrs],
                          O._ptr(dm), *[O._ptr(a) for a in cl])
    finally:
        os.chdir(cwd)
    out = {}
    for name in ("EnergyMomentum", "DotVZ", "Concentration", "ShellConcentration_Catalytic"):
        with open(os.path.join(tmp_path, name + ".dat")) as f:
            text = f.read()
        out[name] = text
    return out


def _nums(text):
    return [float(v) for v in text.split()]


def _cells(seed):
    rng = np.random.default_rng(seed)
    return dict(tF=rng.integers(0, 8, VOL), tG=rng.integers(0, 5, VOL), tI=rng.integers(0, 3, VOL),
                aff=(rng.random(VOL) < 0.02).astype(np.int64))


MASSES = {0: 1.0, 1: 1.25, 3: 0.8}
MSPH = (33.51032163829112, 268.082573106329)


@pytest.mark.parametrize("across", [False, True])
def test_selkov_observables_are_the_reference_output(tmp_path, across):
    st = _state(11 + across, dimer_across=across)
    tpot = 3.25
    ref = _ref_output(tmp_path, st, MASSES, MSPH, tpot, _cells(4))
    em = _nums(ref["EnergyMomentum"])
    dv = _nums(ref["DotVZ"])
    r = R.selkov_observables(st, MASSES, MSPH, [40.0] * 3, tpot)
    assert em[0] == 10 and dv[0] == 10
    assert list(r["ptot"]) == em[1:4]
    assert r["ke_dim"] + r["ke_sol"] == em[4]
    assert em[5] == tpot and r["E_tot"] == em[6]
    assert r["T_sol"] == em[7] and r["T_dim"] == em[8]
    assert r["v_prop"] == dv[1] and list(r["axis"]) == dv[2:5]


def test_selkov_records_are_the_reference_files(tmp_path):
    """selkov_files writes the same bytes as the reference's FORMAT records."""
    from main_amd import selkov_files as F
    try:
        F.format_record(F.FMT23, [1])
    except Exception as exc:   # the engine library is not built here
        pytest.skip(f"libmpcx.so unavailable: {exc}")
    st = _state(5)
    cells = _cells(6)
    tpot = -1.5
    ref = _ref_output(tmp_path, st, MASSES, MSPH, tpot, cells)
    n = st["x"].size
    s = dict(st, NT=n, tpotential=tpot)
    p = dict(box=np.array([40.0, 40.0, 40.0]), Mf=MASSES[0], Mg=MASSES[1], Mi=MASSES[3], Mc=MSPH[0], Mn=MSPH[1])
    c = dict(tF=cells["tF"], tG=cells["tG"], tI=cells["tI"], affected=cells["aff"] == 1)
    assert F.format_record(F.FMT17, [10, *F.energy_momentum(s, p)]) + "\n" == ref["EnergyMomentum"]
    assert F.format_record(F.FMT18, [10, *F.dot_vz(s, p)]) + "\n" == ref["DotVZ"]
    assert F.format_record(F.FMT19, [10, *F.concentration(s, c, p)]) + "\n" == ref["Concentration"]


def test_shells_are_the_reference_shells(tmp_path):
    st = _state(21, n=60000)
    ref = _ref_output(tmp_path, st, MASSES, MSPH, 0.0, _cells(2))
    conc = np.array(_nums(ref["ShellConcentration_Catalytic"])[1:]).reshape(8, 3).T   # (species, shell)
    vols = np.array(R.shell_volumes())
    counts = np.rint(conc * vols).astype(np.int64)
    assert np.array_equal(counts / vols, conc)   # Tot/ShellVol exactly, so the counts are recovered
    got = R.shells(st, [40.0] * 3, st["dm"][0, 0:3])
    assert np.array_equal(got, counts)
    assert got.sum() > 1000


def test_collective_solvent_terms_are_the_reference_expressions(tmp_path):
    """output_mod.f95:696-738 against Output_SelkovDimer.f90:250-283 on the
    same particles (A, B, S as F, G, I; dimer at rest, so sumPX = sumPXsol and
    Ekin = KEsol)."""
    st = _state(31, flags=(0, 1, 2))
    sk = dict(st, flag=np.where(st["flag"] == 2, 3, st["flag"]))
    masses_co = {0: MASSES[0], 1: MASSES[1], 2: MASSES[3]}
    st["vd"][:] = 0.0
    sk["dm"] = st["dm"].copy(order="F")
    sk["dm"][:, 3:6] = 0.0
    ref = _ref_output(tmp_path, sk, MASSES, MSPH, 0.0, _cells(3))
    em = _nums(ref["EnergyMomentum"])
    r = R.observables(st, masses_co, {0: 0, 1: 1, 2: 2}, MSPH, 1, [40.0] * 3, False, 0.0)
    assert list(r["psol"]) == em[1:4]
    assert r["ke_sol"] == em[4]
    assert r["T_sol"] == em[7]


def test_collective_dimer_terms_against_exact_arithmetic():
    """The per-sphere momentum and kinetic energy of output_mod.f95:719-732
    against exact rationals: right masses on the right spheres, the factor
    1/2, the sums over every dimer (a dropped term or swapped mass is off by
    far more than the few ulps allowed)."""
    rng = np.random.default_rng(3)
    nd = 5
    vd = rng.normal(0, 0.1, (2 * nd, 3))
    xd = rng.uniform(0, 30, (2 * nd, 3))
    mc, mn = 33.5, 268.1
    st = dict(x=np.zeros(1), y=np.zeros(1), z=np.zeros(1), vx=np.zeros(1), vy=np.zeros(1), vz=np.zeros(1),
              flag=np.zeros(1, np.int64), xd=xd, vd=vd)
    r = R.observables(st, {0: 1.0}, {0: 0}, (mc, mn), nd, [30.0] * 3, False, 0.0)
    F = Fraction
    ke = sum(F(1, 2) * F(m) * sum(F(v) ** 2 for v in vd[s]) for k in range(nd)
             for s, m in ((k, mc), (k + nd, mn)))
    assert abs(F(r["ke_dim"]) - ke) <= 8 * F(np.spacing(float(ke)))
    for a in range(3):
        p = sum(F(mc) * F(vd[k][a]) + F(mn) * F(vd[k + nd][a]) for k in range(nd))
        assert abs(F(float(r["ptot"][a])) - p) <= 16 * F(np.spacing(float(abs(p)) + 1e-300) + 1e-17)
    assert r["T_dim"] == r["ke_dim"] * float(np.float32(2) / np.float32(3))
