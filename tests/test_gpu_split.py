"""Split-precision storage against the FP64 engine (device Philox draws, same
key, same initial state): after one interval nearly every particle agrees to
FP32 resolution; over a long run the temperature and B-particle count track
the FP64 run, the total momentum (not conserved: the walls reverse every
velocity component) keeps the FP64 run's magnitude, and the dimer's mean
propulsion velocity agrees within 2 sigma (block-averaged standard errors of
both runs; every run here is deterministic, so the outcome is fixed by the
seeds).  The errors are printed per quantity.  The cell sums are exact
integers, so the far flight beside the near engine and the serial order give
bitwise the same states.  The 100-dimer slab also runs in split storage."""
from __future__ import annotations

import numpy as np
import pytest

from main_amd import systems as S
from main_amd.collective import CollectiveEngine
from main_amd.tour import cell_tour

pytestmark = pytest.mark.gpu


def _engine(p, s, split, seed=77):
    e = CollectiveEngine(p, rng_mode=1, seed=seed, split=split)
    e.set_state(s, tour=cell_tour(*[int(v) for v in p["box"]]))
    e.setup(do_sort=True)
    return e


def _params(L=30, **kw):
    base = dict(Lx=float(L), Ly=float(L), Lz=float(L), numdims=1, rhoa=10.0, Rc=2.0, Rn=4.0, Dcn=6.5, EbN=0.1,
                T=1.0 / 6.0, skinrad=0.3, MDt=0.01, MPCt=0.5, sort_step=50, k1=1e-3)
    base.update(kw)
    return S.collective_params(**base)


def test_split_one_interval_matches_fp64_to_fp32_resolution():
    p = _params()
    s = S.collective_initial_state(p, seed=5)
    out = []
    for split in (False, True):
        e = _engine(p, s, split)
        try:
            e.run(1)
            out.append(e.get_state())
        finally:
            e.close()
    a, b = out
    dx = np.abs(np.stack([b[k] - a[k] for k in ("x", "y", "z")]))
    dx = np.minimum(dx, np.abs(dx - p["box"][:, None]))
    dv = np.abs(np.stack([b[k] - a[k] for k in ("vx", "vy", "vz")]))
    frac_x = np.mean(np.any(dx > 1e-4, axis=0))
    frac_v = np.mean(np.any(dv > 1e-4, axis=0))
    print(f"one interval: median |dx| {np.median(dx):.2e}, median |dv| {np.median(dv):.2e}, "
          f"particles off by >1e-4: x {frac_x:.2e}, v {frac_v:.2e}")
    assert np.median(dx) < 1e-6 and np.median(dv) < 1e-6
    assert frac_x < 1e-3 and frac_v < 2e-3      # cell-boundary ties of the shifted grid
    assert np.array_equal(a["flag"] == 1, b["flag"] == 1) or np.mean(a["flag"] != b["flag"]) < 1e-4
    assert abs(a["xd"][0, 0] - b["xd"][0, 0]) < 1e-4


def test_split_overlap_is_bitwise_the_serial_order():
    """Far flight and binning beside the near engine (second stream) against
    the same kernels one after the other: identical states, bit for bit,
    through a re-sort (exact integer cell sums, no order dependence)."""
    p = _params(L=24, sort_step=100)   # re-sort after the 2nd collision
    s = S.collective_initial_state(p, seed=12)
    out = []
    for ov in (True, False):
        e = CollectiveEngine(p, rng_mode=1, seed=5, split=True, overlap=ov)
        e.set_state(s, tour=cell_tour(24, 24, 24))
        e.setup(do_sort=True)
        try:
            e.run(4)
            out.append(e.get_state())
        finally:
            e.close()
    a, b = out
    assert a["cstep"] == b["cstep"] == 4 * p["MPCtstep"]
    for k in ("x", "y", "z", "vx", "vy", "vz", "flag", "xd", "yd", "zd", "vxd", "fxd"):
        assert np.array_equal(a[k], b[k]), k


def _series(e, n, p):
    vs, ts, ps, nb = [], [], [], []
    for _ in range(n):
        e.run(1)
        o = e.observables()
        vs.append(o["v_prop"])
        ts.append(o["T_sol"])
        ps.append(np.sqrt(o["p_x"] ** 2 + o["p_y"] ** 2 + o["p_z"] ** 2))
        nb.append(o["nB"])
    return np.array(vs), np.array(ts), np.array(ps), np.array(nb)


def _block_se(x, nb=20):
    b = x[: len(x) // nb * nb].reshape(nb, -1).mean(axis=1)
    return b.std(ddof=1) / np.sqrt(nb)


def test_split_long_run_statistics_track_fp64():
    p = _params(L=24)
    s = S.collective_initial_state(p, seed=6)
    res = {}
    for split in (False, True):
        e = _engine(p, s, split, seed=91)
        try:
            e.run(100)
            res[split] = _series(e, 1500, p)
        finally:
            e.close()
    (v0, t0, p0, b0), (v1, t1, p1, b1) = res[False], res[True]
    se = np.hypot(_block_se(v0), _block_se(v1))
    print(f"v_prop: fp64 {v0.mean():.5f}, split {v1.mean():.5f}, 2 sigma {2 * se:.5f}")
    print(f"T_sol rel diff {abs(t1.mean() / t0.mean() - 1):.2e}; |P| fp64 {p0.max():.2e} split {p1.max():.2e}; "
          f"nB rel diff {abs(b1.mean() / b0.mean() - 1):.2e}")
    assert abs(v1.mean() - v0.mean()) < 2.0 * se
    assert abs(t1.mean() / t0.mean() - 1.0) < 2e-3
    assert abs(b1.mean() / b0.mean() - 1.0) < 0.03
    r = np.sqrt(np.mean(p1 ** 2) / np.mean(p0 ** 2))
    print(f"RMS |P| split / fp64 {r:.3f}")
    assert 0.5 < r < 2.0


def test_split_many_dimer_slab_runs():
    p = S.collective_params(Lx=200.0, Ly=200.0, Lz=10.0, numdims=100, rhoa=10.0, sort_step=500)
    d = S.collective_dimers(p, seed=3)
    e = CollectiveEngine(p, rng_mode=1, seed=3, split=True)
    try:
        e.set_dimers(d, tour=cell_tour(200, 200, 10))
        e.init_solvent(3)
        e.setup(do_sort=True)
        e.run(3)
        g = e.get_state()
    finally:
        e.close()
    assert g["cstep"] == 3 * p["MPCtstep"]
    ke = 0.5 * np.mean(g["vx"] ** 2 + g["vy"] ** 2 + g["vz"] ** 2)
    assert abs(ke / (1.5 * p["T"]) - 1.0) < 0.05


def _slab(split, seed=3):
    p = S.collective_params(Lx=200.0, Ly=200.0, Lz=10.0, numdims=100, rhoa=10.0, sort_step=500)
    e = CollectiveEngine(p, rng_mode=1, seed=seed, split=split)
    e.set_dimers(S.collective_dimers(p, seed=seed), tour=cell_tour(200, 200, 10))
    e.init_solvent(seed)
    e.setup(do_sort=True)
    return p, e


def _body_b(e, nd, nr=12, nz=24):
    """B counts in the body frames of all dimers, summed: (nz, nr)."""
    return sum(e.body_density(nr=nr, nz=nz, rmax=12.0, zmin=-16.0, zmax=12.0, dimer=k)[1].astype(np.float64)
               for k in range(nd))


def test_split_slab_errors_against_fp64():
    """The 100-dimer slab (200 x 200 x 10 cells, 4.0e6 solvent) in split and
    FP64 storage from the same initial state and Philox key, with the errors
    of split storage stated per quantity: the cell momenta after one interval
    (cells whose membership differs, by cell-boundary ties of FP32 offsets,
    counted apart), then over 60 collisions the temperature, the B count, the
    motor velocity (mean over the 100 dimers, within 2 sigma of block
    averages) and the body-frame B-density field summed over the dimers
    (against its counting noise)."""
    cells, series, fields = {}, {}, {}
    for split in (False, True):
        p, e = _slab(split)
        try:
            e.run(1)
            cells[split] = e.get_cells()
            v, t, b, f = [], [], [], np.zeros((24, 12))
            for c in range(60):
                e.run(1)
                o = e.observables()
                v.append(o["v_prop"]); t.append(o["T_sol"]); b.append(o["nB"])
                if c >= 30 and c % 3 == 0:
                    f += _body_b(e, 100)
            series[split] = (np.array(v), np.array(t), np.array(b))
            fields[split] = f
        finally:
            e.close()
    a, c = cells[False], cells[True]
    same = a["tcs"] == c["tcs"]
    occ = same & (a["tcs"] > 0)
    dv = np.abs(np.stack([c[k][occ] - a[k][occ] for k in ("vxcm", "vycm", "vzcm")]))
    print(f"cell momenta after one interval: {np.mean(~same):.2e} of the cells differ in membership; over the "
          f"others |d vcm| median {np.median(dv):.2e}, max {dv.max():.2e}")
    assert np.mean(~same) < 1e-3
    assert np.median(dv) < 1e-6 and dv.max() < 1e-4
    (v0, t0, b0), (v1, t1, b1) = series[False], series[True]
    se = np.hypot(_block_se(v0, 10), _block_se(v1, 10))
    print(f"slab motor velocity: fp64 {v0.mean():.5f}, split {v1.mean():.5f}, 2 sigma {2 * se:.5f}; T_sol rel diff "
          f"{abs(t1.mean() / t0.mean() - 1):.2e}; nB rel diff {abs(b1.mean() / b0.mean() - 1):.2e}")
    assert abs(v1.mean() - v0.mean()) < 2.0 * se
    assert abs(t1.mean() / t0.mean() - 1.0) < 2e-3
    assert abs(b1.mean() / b0.mean() - 1.0) < 0.03
    f0, f1 = fields[False], fields[True]
    m = (f0 + f1) > 20
    chi = np.mean((f1[m] - f0[m]) ** 2 / (f0[m] + f1[m]))
    rel = np.linalg.norm(f1 - f0) / np.linalg.norm(f0)
    print(f"body-frame B field: relative L2 difference {rel:.3f}, (diff^2 / counting variance) {chi:.2f} over "
          f"{int(m.sum())} bins")
    assert chi < 3.0


def test_split_keeps_particles_below_the_top_wall():
    """FP32 offsets cannot hold a z within half an ulp below Lz: a far particle
    that ends its flight there (or converts from FP64 there) must stay in the
    top cell instead of leaving the cell arrays (a regression test: such a
    particle used to get cell z index Lz and stop the run with a range error)."""
    p = _params(L=20)
    s = S.collective_initial_state(p, seed=8)
    Lz = float(p["box"][2])
    cases = [(Lz - 2.0 ** -24, 6e-8), (Lz - 1e-12, 0.0), (Lz - 2.0 ** -24, 2.0 ** -26), (Lz - 2.0 ** -25, 1e-9)]
    for n, (z, vz) in enumerate(cases):
        s["x"][n], s["y"][n], s["z"][n] = 1.5 + 0.1 * n, 1.5, z
        s["vx"][n], s["vy"][n], s["vz"][n] = 0.0, 0.0, vz
    e = _engine(p, s, True)
    try:
        e.run(2)
        g = e.get_state()
    finally:
        e.close()
    assert np.all(g["z"] < Lz) and np.all(g["z"] >= 0.0)


def test_split_stores_x_equal_to_the_box_edge_as_its_periodic_image():
    """x = Lx (y = Ly) is the periodic image of 0: the reference's wrap
    "if (x < 0) x = x + L" yields it when a flight ends a rounding error below
    0, and a near particle whose FP32 start and velocity put its flight's end
    on the boundary exactly (x + v tau = 0) reaches it with either sign of the
    FP64 steps' rounding.  Split storage keeps it as cell 0, offset 0, where
    the conversion from FP64 used to reject it and stop the run with a range
    error (a regression test: a replica of the 256-system ensemble stopped so
    in its 94th collision).  The set-up without a re-sort stores every particle
    through that same conversion."""
    p = _params(L=20)
    s = S.collective_initial_state(p, seed=8)
    L = float(p["box"][0])
    sph = np.concatenate([np.stack([s["xd"][:, i], s["yd"][:, i], s["zd"][:, i]], axis=1) for i in (0, 1)])
    cases = [(L, 12.5, 4.5), (12.5, L, 3.5), (L, L, 16.5)]   # FP32-exact coordinates
    for x, y, z in cases:
        for c in sph:
            e = np.array([x, y, z]) - c
            e[:2] -= L * np.round(e[:2] / L)
            assert np.linalg.norm(e) > 6.0, "the planted particles stay clear of the spheres"
    picks = [0, 1, 2]
    for n, (x, y, z) in zip(picks, cases):
        s["x"][n], s["y"][n], s["z"][n] = x, y, z
    e = CollectiveEngine(p, rng_mode=1, seed=77, split=True)
    try:
        e.set_state(s, tour=cell_tour(20, 20, 20))
        e.setup(do_sort=False)
        g = e.get_state()
    finally:
        e.close()
    for n, (x, y, z) in zip(picks, cases):
        assert (g["x"][n], g["y"][n], g["z"][n]) == (x % L, y % L, z), n


def test_split_reports_a_particle_thrown_out_of_the_box():
    """A near particle that the FP64 near engine carries far outside the box
    (planted here with vx = -1e25) is reported when the near set goes back to
    split storage, and is never binned (a regression test: its cell
    coordinates used to wrap into a negative collision-cell index)."""
    from main_amd._lib import MpcxError
    p = _params(L=20)
    s = S.collective_initial_state(p, seed=8)
    pos = np.stack([s["x"], s["y"], s["z"]], axis=1)
    sph = np.concatenate([np.stack([s["xd"][:, i], s["yd"][:, i], s["zd"][:, i]], axis=1) for i in (0, 1)])
    d = np.min(np.linalg.norm(pos[:, None, :] - sph[None, :, :], axis=2), axis=1)
    pick = int(np.argmax(d))   # the particle farthest from both spheres
    s["vx"][pick] = -1e25
    e = _engine(p, s, True)
    try:
        with pytest.raises(MpcxError, match=r"outside the cell arrays \[sb_near .*particle \d+"):
            e.run(1)
    finally:
        e.close()


def test_split_run_after_a_failed_run_equals_a_fresh_engine():
    """A run stopped by an error leaves partial sums in the accumulator
    copies; the next run clears them first (runs that end cleanly leave them
    cleared), so a new state run in the same engine matches a fresh engine."""
    from main_amd._lib import MpcxError
    p = _params(L=20)
    good = S.collective_initial_state(p, seed=9)
    bad = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in good.items()}
    pos = np.stack([bad["x"], bad["y"], bad["z"]], axis=1)
    sph = np.concatenate([np.stack([bad["xd"][:, i], bad["yd"][:, i], bad["zd"][:, i]], axis=1) for i in (0, 1)])
    pick = int(np.argmax(np.min(np.linalg.norm(pos[:, None, :] - sph[None, :, :], axis=2), axis=1)))
    bad["vx"][pick] = -1e25
    e = _engine(p, bad, True, seed=31)
    try:
        with pytest.raises(MpcxError):
            e.run(1)
        e.set_state(good, tour=cell_tour(20, 20, 20))
        e.setup(do_sort=True)
        e.run(3)
        a = e.get_state()
    finally:
        e.close()
    f = _engine(p, good, True, seed=31)
    try:
        f.run(3)
        b = f.get_state()
    finally:
        f.close()
    for k in ("x", "y", "z", "vx", "vy", "vz", "flag", "xd", "vxd"):
        assert np.array_equal(a[k], b[k]), k
