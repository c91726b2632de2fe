"""GPU parity and properties of the device-resident collective engine.

Parity (host Mersenne-Twister draws, the reference's draw order): the engine
and the oracle advance the same seeded system through whole collision
intervals; integer state (flags, cell indices, census and collision-cell
counts, inside list, neighbour counts, cstep/time_count/time_penetrate) must
be identical and positions/velocities within 1e-12 relative
(tests/parity_cases.py states the scales).  The golden replays compare the
engine's integer arrays with the SHA-256 the reference's own Fortran modules
produced (tests/golden/reference_runs.json).

Properties at the full 100^3 size (device Philox draws, where no oracle run is
affordable): particle number, confinement, temperature, and bit-identical
reruns.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import sys

import numpy as np
import pytest

from main_amd import systems as S
from main_amd._lib import MtmodStream, device_synchronize
from main_amd.collective import CollectiveEngine, host_draws
from main_amd.tour import cell_tour
from oracle import oracle as O
from tests.parity_cases import collective_parity_case, small_collective_params

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _reactive_grid(p, kind):
    rb = np.zeros(int(p["box"][0]) * int(p["box"][1]), np.int32)
    O.oracle_lib().or_co_reactive_grid_setup(O._ptr(p["box"]), O._ptr(rb), kind)
    return rb


CASES = {
    "basic_3coll": dict(ncoll=3, seed=3),
    "basic_3coll_event_replay": dict(ncoll=3, seed=3, overlap=False),
    "sort_and_species": dict(ncoll=4, seed=4, flags_mix=True, p=("small", dict(sort_step=20)), tour="hilbert"),
    "sort_and_species_event_replay": dict(ncoll=4, seed=4, flags_mix=True, p=("small", dict(sort_step=20)),
                                          tour="hilbert", overlap=False),
    "two_dimers_sort": dict(ncoll=3, seed=7, flags_mix=True,
                            p=("small", dict(numdims=2, Lx=20.0, Ly=20.0, Lz=24.0, sort_step=20, EbN=0.1)),
                            tour="hilbert"),
    "three_dimers_sort": dict(ncoll=3, seed=11, flags_mix=True,
                              p=("small", dict(numdims=3, Lx=18.0, Ly=18.0, Lz=16.0, sort_step=20, EbN=0.1)),
                              tour="hilbert"),
    "four_dimers": dict(ncoll=3, seed=12, p=("small", dict(numdims=4, Lx=18.0, Ly=18.0, Lz=14.0, EbN=0.3))),
    "four_dimers_one_workgroup": dict(ncoll=3, seed=12, near_mode=0,
                                      p=("small", dict(numdims=4, Lx=18.0, Ly=18.0, Lz=14.0, EbN=0.3))),
    "two_dimers_clusters": dict(ncoll=3, seed=7, flags_mix=True, near_mode=1,
                                p=("small", dict(numdims=2, Lx=20.0, Ly=20.0, Lz=24.0, sort_step=20, EbN=0.1)),
                                tour="hilbert"),
    "eight_dimers_clusters_sort": dict(ncoll=4, seed=13, flags_mix=True, near_mode=1,
                                       p=("small", dict(numdims=8, Lx=28.0, Ly=28.0, Lz=16.0, sort_step=20)),
                                       tour="hilbert"),
    "eight_dimers_global_sort": dict(ncoll=4, seed=13, flags_mix=True,
                                     p=("small", dict(numdims=8, Lx=28.0, Ly=28.0, Lz=16.0, sort_step=20)),
                                     tour="hilbert"),
    "two_dimers_global": dict(ncoll=3, seed=7, flags_mix=True, near_mode=2,
                              p=("small", dict(numdims=2, Lx=20.0, Ly=20.0, Lz=24.0, sort_step=20, EbN=0.1)),
                              tour="hilbert"),
    "twenty_dimers_dense": dict(ncoll=3, seed=14, p=("small", dict(numdims=20, Lx=40.0, Ly=40.0, Lz=16.0, rhoa=4.0))),
    "reactive_wall": dict(ncoll=3, seed=8, p=("small", dict(k1=0.2, Lz=12.0)), react="wall2"),
    "dense_skin05": dict(ncoll=2, seed=9, flags_mix=True, p=("small", dict(rhoa=8.0, skinrad=0.5, Rn=2.5))),
}


@pytest.mark.parametrize("name", list(CASES))
def test_engine_matches_oracle(name):
    kw = dict(CASES[name])
    _say(f"[parity] {name}")
    if "p" in kw:
        kw["p"] = small_collective_params(**kw["p"][1])
    p = kw.get("p") or small_collective_params()
    if kw.pop("tour", None) == "hilbert":
        kw["tour"] = cell_tour(int(p["box"][0]), int(p["box"][1]), int(p["box"][2]))
    if kw.pop("react", None) == "wall2":
        kw["react_box"] = _reactive_grid(p, 2)
    r = collective_parity_case(**kw)
    _say(f"[parity] {name}: {r['summary']}")
    assert r["ok"], r["summary"]


def test_engine_matches_oracle_30cube_headline_physics():
    """The reference's CPU-runnable size: 30^3 cells, 10 per cell (266,984
    solvent), Rc 2 / Rn 4 / bond 6.5, tau 0.5 = 50 x 0.01, Hilbert sort."""
    from bench import headline_params
    p = headline_params(30)
    p["sort_step"] = 100     # a re-sort inside the compared window
    _say("[parity] 30^3 headline physics, 3 collisions")
    r = collective_parity_case(ncoll=3, seed=21, p=p, tour=cell_tour(30, 30, 30))
    _say(f"[parity] 30^3: {r['summary']}")
    assert r["ok"], r["summary"]


def _h(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("case,near_mode,overlap", [
    ("co_one_dimer", None, True), ("co_one_dimer", None, False), ("co_two_dimers_sort_wall", None, True),
    ("co_two_dimers_sort_wall", None, False), ("co_six_dimers_sort", 1, True), ("co_two_dimers_sort_wall", 1, True),
    ("co_two_dimers_sort_wall", 2, True), ("co_six_dimers_sort", None, True)])
def test_engine_integer_state_equals_reference_fixture(golden, case, near_mode, overlap):
    """Flags and cell arrays hashed from the reference's Fortran run, in
    every near-engine mode (0 one workgroup, 1 clusters, 2 cooperative
    workgroups over the near set; None the default for the dimer count), with
    the far particles' single flight beside the near engine (overlap, used by
    the one-workgroup mode) and with the event replay after it."""
    from tests.golden.make_golden import collective_inputs
    rec = golden["cases"][case]
    cs = rec["case"]
    p, s = collective_inputs(cs)
    L = [int(v) for v in p["box"]]
    tour = cell_tour(*L) if cs["sort"] else np.arange(1, p["vol"] + 1, dtype=np.int64)
    ncoll = cs["steps"] // p["MPCtstep"]
    assert ncoll * p["MPCtstep"] == cs["steps"]
    draws = host_draws(MtmodStream(mt=s["mt"], mti=s["mti"]), p, ncoll)
    eng = CollectiveEngine(p, rng_mode=0, near_mode=near_mode, overlap=overlap)
    try:
        eng.set_state(s, tour=tour)
        eng.setup(do_sort=cs["sort"])
        eng.run(ncoll, draws)
        g = eng.get_state()
        cells = eng.get_cells()
        lists = eng.get_lists()
    finally:
        eng.close()
    got = {"flag": _h(g["flag"].astype(np.int32)), "ncn": _h(cells["ncn"].astype(np.int64)),
           "tcs": _h(cells["tcs"]), "tA": _h(cells["tA"][: p["vol"]]), "tB": _h(cells["tB"][: p["vol"]]),
           "tS": _h(cells["tS"][: p["vol"]]), "nbnum": _h(lists["nbnum"].astype(np.int64))}
    bad = [k for k, v in got.items() if v != rec["arrays"][k]]
    assert not bad, f"{case}: differs from the reference run in {bad}"
    for k in ("cstep", "time_count", "time_penetrate"):
        assert g[k] == rec["scalars"][k], k
    # the inside/outside split (the fixture's il/ol arrays hold earlier
    # builds' entries past the counts, so the counts are what compare)
    assert lists["in_list"].size == rec["scalars"]["in_num"]
    assert lists["out_list"].size == rec["scalars"]["out_num"]


def _full_size_engine(L, seed):
    from bench import headline_params
    p = headline_params(L)
    s = S.collective_initial_state(p, seed=seed)
    eng = CollectiveEngine(p, rng_mode=1, seed=seed)
    eng.set_state(s, tour=cell_tour(L, L, L))
    eng.setup(do_sort=True)
    return p, s, eng


def test_full_size_conservation_and_confinement():
    """100^3 cells (9,996,984 solvent), 12 collisions with a re-sort, Philox draws.
    (The z walls bounce particles back, so no momentum component is conserved.)"""
    L = 100
    _say("[props] building 100^3")
    p, s, eng = _full_size_engine(L, 4357787)
    try:
        m = p["Ma"]
        eng.run(12)
        device_synchronize()
        g = eng.get_state()
        cells = eng.get_cells()
    finally:
        eng.close()
    _say(f"[props] cstep {g['cstep']}")
    assert g["cstep"] == 12 * p["MPCtstep"]
    assert set(np.unique(g["flag"])) <= {0, 1}
    assert np.all((g["x"] >= 0) & (g["x"] < L) & (g["y"] >= 0) & (g["y"] < L))
    assert np.all((g["z"] >= 0) & (g["z"] <= L))
    # every particle lies in exactly one collision cell
    assert int(cells["tcs"].sum()) == p["NT"]
    assert np.all((cells["ncn"] >= 1) & (cells["ncn"] <= p["mpcnumcell"]))
    # the census counts only particles outside the sphere cutoffs
    assert 0 < int(cells["tA"].sum() + cells["tB"].sum()) <= p["NT"]
    # equipartition (2/3) KE / N = kT, within the 3/sqrt(N)-level noise
    ke = 0.5 * m * (g["vx"] ** 2 + g["vy"] ** 2 + g["vz"] ** 2).sum()
    T = (2.0 / 3.0) * ke / p["NT"]
    assert abs(T / p["T"] - 1.0) < 2e-3, T


def test_reruns_are_bit_identical():
    """Same seed, same state -> identical arrays (ordered sums everywhere)."""
    outs = []
    for rep in range(2):
        p, s, eng = _full_size_engine(40, 99)
        try:
            eng.run(5)
            outs.append(eng.get_state())
        finally:
            eng.close()
    for k in ("x", "y", "z", "vx", "vy", "vz", "flag", "xd", "vxd"):
        assert np.array_equal(outs[0][k], outs[1][k]), k


def test_global_engine_reruns_are_bit_identical():
    """The cooperative near engine over many workgroups (the reference input's
    20 dimers in 64x64x32, 1.25e6 particles): two runs from the same seed end
    bitwise equal.  Its MD-step barriers skip the L2 write-back (the sphere
    sums crossing them are agent-scope atomics and no entry changes owner
    between list rebuilds), so a missed exchange through plain memory would
    show here as a difference between reruns."""
    from bench import MANY_DIMER
    p = S.collective_params(**MANY_DIMER["reference_input_20_dimers_64x64x32"], rhoa=10.0, Rc=2.0, Rn=4.0, Dcn=6.5,
                            EbN=0.1, T=1.0 / 6.0, skinrad=0.3, MDt=0.01, MPCt=0.5, sort_step=500, k1=1e-3)
    out = []
    for _ in range(2):
        eng = CollectiveEngine(p, rng_mode=1, seed=77)
        try:
            eng.set_dimers(S.collective_dimers(p, seed=77), tour=cell_tour(64, 64, 32))
            eng.init_solvent(77)
            eng.setup(do_sort=True)
            eng.run(12)
            out.append(eng.get_state())
        finally:
            eng.close()
    a, b = out
    for k in ("x", "y", "z", "vx", "vy", "vz", "flag", "xd", "yd", "zd", "vxd", "vyd", "vzd", "fxd"):
        assert np.array_equal(a[k], b[k]), k
    assert a["cstep"] == b["cstep"] == 12 * p["MPCtstep"]


def test_overlap_matches_the_event_replay():
    """The far particles' single flight beside the near engine against the
    bitwise replay of the rebuild events after it, 40^3 headline physics,
    Philox draws, re-sorts every 2 collisions.  Per interval, from one
    bit-identical state (k replayed intervals, then one of each kind):
    integer state identical, positions and velocities within 1e-12 -- the
    flights differ by rounding only.  Over 4 intervals in a row the integer
    state stays identical; the floating-point state is not compared there,
    since the dimer's LJ dynamics amplify last-bit differences (one particle
    reached 1.2e-12 after 4 collisions)."""
    from bench import headline_params
    from tests.parity_cases import close
    p = headline_params(40)
    p["sort_step"] = 100
    s = S.collective_initial_state(p, seed=31)
    tour = cell_tour(40, 40, 40)
    vscale = np.sqrt(p["T"] / p["Ma"])

    def engine(overlap):
        e = CollectiveEngine(p, rng_mode=1, seed=31, overlap=overlap)
        e.set_state(s, tour=tour)
        e.setup(do_sort=True)
        return e

    def same_integers(g, c, r, rc):
        assert np.array_equal(g["flag"], r["flag"])
        for k in ("ncn", "tcs", "tA", "tB", "tS"):
            assert np.array_equal(c[k], rc[k]), k
        for k in ("cstep", "time_count", "time_penetrate"):
            assert g[k] == r[k], k

    differs = 0
    ref = engine(False)
    try:
        for k in range(4):   # the (k+1)-th interval; the 2nd and 4th re-sort
            ref.run(1)
            r, rc = ref.get_state(), ref.get_cells()
            a = engine(False)
            try:
                if k:
                    a.run(k)
                a.set_overlap(True)
                a.run(1)
                g, c = a.get_state(), a.get_cells()
            finally:
                a.close()
            _say(f"[overlap] interval {k + 1}")
            same_integers(g, c, r, rc)
            for key, sc in (("x", 1.0), ("y", 1.0), ("z", 1.0), ("vx", vscale), ("vy", vscale), ("vz", vscale),
                            ("xd", 1.0), ("yd", 1.0), ("zd", 1.0), ("vxd", 1e-3), ("vyd", 1e-3), ("vzd", 1e-3)):
                n, err = close(g[key], r[key], sc)
                assert n == 0, f"interval {k + 1}, {key}: {n} values beyond 1e-12 (worst {err:.2e})"
            differs += int(not np.array_equal(g["x"], r["x"]))
        a = engine(True)
        try:
            a.run(4)
            same_integers(a.get_state(), a.get_cells(), r, rc)
        finally:
            a.close()
    finally:
        ref.close()
    assert differs > 0   # the single flight did run (its last bits differ from the replay's)


def test_results_do_not_depend_on_the_bulk_grid():
    """co_bulk is a grid-stride loop whose per-cell ranks come from atomics in
    whatever order the workgroups run; the collision kernel orders the members,
    so every grid gives bit-identical states (40^3, 12 collisions with a re-sort)."""
    from bench import headline_params
    from main_amd._lib import check, lib
    p = headline_params(40)
    p["sort_step"] = 250     # a re-sort every 5 collisions
    s = S.collective_initial_state(p, seed=12)
    out = []
    for grid in (0, 64, 1000, 20000):
        eng = CollectiveEngine(p, rng_mode=1, seed=12)
        try:
            check(lib().mpcx_co_set_bulk_grid(eng.h, C.c_int32(grid)), "mpcx_co_set_bulk_grid")
            eng.set_state(s, tour=cell_tour(40, 40, 40))
            eng.setup(do_sort=True)
            eng.run(12)
            out.append(eng.get_state())
        finally:
            eng.close()
    for g in out[1:]:
        for k in ("x", "y", "z", "vx", "vy", "vz", "flag", "xd", "vxd"):
            assert np.array_equal(np.asarray(g[k]).view(np.uint8), np.asarray(out[0][k]).view(np.uint8)), k
