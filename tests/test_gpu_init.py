"""Initial solvent generated on the device (initial_mod.f95:122-345 with
Philox draws): every particle outside every sphere's cutoff and inside the
box, species by index range, Gaussian velocities of variance T/m per species,
zero total momentum, uniform density (per-slab counts within 5 sigma of the
free-volume expectation), reproducible per seed; the 100-dimer 200x200x10
slab and the reference input's 20 dimers in 64x64x32 then advance without a
device error."""
from __future__ import annotations

import numpy as np
import pytest

from main_amd import systems as S
from main_amd.collective import CollectiveEngine
from main_amd.tour import cell_tour

pytestmark = pytest.mark.gpu


def _engine(p, seed):
    d = S.collective_dimers(p, seed=seed)
    e = CollectiveEngine(p, rng_mode=1, seed=seed)
    e.set_dimers(d, tour=cell_tour(*[int(v) for v in p["box"]]))
    e.init_solvent(seed)
    return e, d


def test_device_initial_solvent_statistics():
    p = S.collective_params(Lx=24.0, Ly=24.0, Lz=20.0, numdims=3, rhoa=8.0, rhob=2.0, Ma=1.0, Mb=2.0, Ms=0.5)
    e, d = _engine(p, 17)
    try:
        g = e.get_state()
        e2, _ = _engine(p, 17)
        g2 = e2.get_state()
        e2.close()
    finally:
        e.close()
    NT, box = p["NT"], p["box"]
    x, y, z = g["x"], g["y"], g["z"]
    assert np.all((x >= 0) & (x < box[0]) & (y >= 0) & (y < box[1]) & (z >= 0) & (z < box[2]))
    for k in range(p["numdims"]):
        for i, R in enumerate((p["Rc"], p["Rn"])):
            c = np.array([d["xd"][k, i], d["yd"][k, i], d["zd"][k, i]])
            dx = x - c[0]; dx -= box[0] * np.round(dx / box[0])
            dy = y - c[1]; dy -= box[1] * np.round(dy / box[1])
            dz = z - c[2]
            assert np.min(dx * dx + dy * dy + dz * dz) >= (R * p["cutoff"]) ** 2
    f = g["flag"]
    na, nb = p["Na0"], p["Nb0"]
    assert np.all(f[:na] == 0) and np.all(f[na:na + nb] == 1) and np.all(f[na + nb:] == 2)
    v = np.stack([g["vx"], g["vy"], g["vz"]], axis=1)
    assert np.all(np.abs(v.sum(axis=0)) / NT < 1e-13)
    for sl, m in ((slice(0, na), p["Ma"]), (slice(na, na + nb), p["Mb"])):
        var = v[sl].var(axis=0)
        assert np.all(np.abs(var / (p["T"] / m) - 1.0) < 0.03), var
    # per-slab counts against the free volume of each slab (host Monte Carlo, 4M points)
    counts, _ = np.histogram(x, bins=12, range=(0, box[0]))
    rng = np.random.default_rng(5)
    u = rng.random((4_000_000, 3)) * box
    free = np.ones(len(u), bool)
    for k in range(p["numdims"]):
        for i, R in enumerate((p["Rc"], p["Rn"])):
            c = np.array([d["xd"][k, i], d["yd"][k, i], d["zd"][k, i]])
            dd = u - c
            dd[:, :2] -= box[:2] * np.round(dd[:, :2] / box[:2])
            free &= (dd * dd).sum(axis=1) >= (R * p["cutoff"]) ** 2
    frac, _ = np.histogram(u[free, 0], bins=12, range=(0, box[0]))
    expect = NT * frac / frac.sum()
    assert np.all(np.abs(counts - expect) < 5 * np.sqrt(expect) + 3 * expect * np.sqrt(12.0 / free.sum())), \
        (counts, expect)
    assert np.array_equal(g["x"], g2["x"]) and np.array_equal(g["vz"], g2["vz"])


@pytest.mark.parametrize("name", ["slab_100_dimers", "reference_input_20_dimers"])
def test_many_dimer_systems_start_and_run(name):
    if name == "slab_100_dimers":
        p = S.collective_params(Lx=200.0, Ly=200.0, Lz=10.0, numdims=100, rhoa=10.0, sort_step=500)
    else:
        p = S.collective_params(Lx=64.0, Ly=64.0, Lz=32.0, numdims=20, rhoa=10.0, sort_step=500)
    print(f"[init] {name}: NT={p['NT']}", flush=True)
    e, _ = _engine(p, 4357787)
    try:
        e.setup(do_sort=True)
        e.run(3)
        g = e.get_state()
    finally:
        e.close()
    assert g["cstep"] == 3 * p["MPCtstep"]
    z = g["z"]
    assert np.all((z >= 0) & (z <= p["box"][2]))
    ke = 0.5 * np.mean(g["vx"] ** 2 + g["vy"] ** 2 + g["vz"] ** 2)
    assert abs(ke / (1.5 * p["T"]) - 1.0) < 0.05
