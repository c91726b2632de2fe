"""Checkpoint and restart of the device engine against the reference.

A run is stopped at a collision step, written to disk (frame, timestep.dat,
mt_state.dat, motor_positions.dat), read back into a new engine with the
restart set-up of restart_mod.f95:416-495 (forces evaluated on the device),
and continued with the restored Mersenne-Twister stream.  The integer state
at the end must equal the SHA-256 of the reference's own run with the same
restart (tests/golden, restart cases) and the floating-point state must
match the oracle within 1e-12 relative; the restart evaluation itself is
compared with the oracle's (dimer forces, pair forces, potential, species).
"""
from __future__ import annotations

import hashlib
import sys

import numpy as np
import pytest

from main_amd import restart as RS
from main_amd._lib import MpcxError, MtmodStream
from main_amd.collective import CollectiveEngine, host_draws
from main_amd.tour import cell_tour
from oracle import oracle as O
from tests.parity_cases import close

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _h(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _inputs(case_name, golden):
    from tests.golden.make_golden import collective_inputs
    cs = golden["cases"][case_name]["case"]
    p, s = collective_inputs(cs)
    L = [int(v) for v in p["box"]]
    tour = cell_tour(*L) if cs["sort"] else np.arange(1, p["vol"] + 1, dtype=np.int64)
    return cs, p, s, tour


@pytest.mark.parametrize("case,near_mode", [("co_one_dimer_restart", None), ("co_two_dimers_restart_sort", None),
                                            ("co_two_dimers_restart_sort", 2), ("co_six_dimers_restart", None)])
def test_restart_through_files_equals_reference(golden, tmp_path, case, near_mode):
    from tests.golden.make_golden import run_collective
    _say(f"[restart] {case} near_mode={near_mode}")
    cs, p, s, tour = _inputs(case, golden)
    ra, steps, S = cs["restart_at"], cs["steps"], p["MPCtstep"]
    assert ra % S == 0 and steps % S == 0
    stream = MtmodStream(mt=s["mt"], mti=s["mti"])
    a = CollectiveEngine(p, rng_mode=0, near_mode=near_mode)
    try:
        a.set_state(s, tour=tour)
        a.setup(do_sort=cs["sort"])
        a.run(ra // S, host_draws(stream, p, ra // S))
        RS.save_checkpoint(a, str(tmp_path), streams=[stream], cm=RS.motor_cm(s))
    finally:
        a.close()
    b = CollectiveEngine(p, rng_mode=0, near_mode=near_mode)
    try:
        r = RS.restart(b, str(tmp_path), s["react_box"], tour)
        assert r["cstep"] == ra and len(r["streams"]) == 1
        assert np.array_equal(r["motor_cm"], RS.motor_cm(s))
        b.run((steps - ra) // S, host_draws(r["streams"][0], p, (steps - ra) // S))
        g = b.get_state()
        cells = b.get_cells()
    finally:
        b.close()
    rec = golden["cases"][case]
    got = {"flag": _h(g["flag"].astype(np.int32)), "ncn": _h(cells["ncn"].astype(np.int64)),
           "tcs": _h(cells["tcs"]), "tA": _h(cells["tA"][: p["vol"]]), "tB": _h(cells["tB"][: p["vol"]]),
           "tS": _h(cells["tS"][: p["vol"]])}
    bad = [k for k, v in got.items() if v != rec["arrays"][k]]
    assert not bad, f"{case}: differs from the reference's restarted run in {bad}"
    for k in ("cstep", "time_count", "time_penetrate"):
        assert g[k] == rec["scalars"][k], k
    ref = run_collective(cs, use_ref=False)
    vscale = np.sqrt(p["T"] / p["Ma"])
    for k, sc in (("x", 1.0), ("y", 1.0), ("z", 1.0), ("vx", vscale), ("vy", vscale), ("vz", vscale),
                  ("xd", 1.0), ("yd", 1.0), ("zd", 1.0), ("vxd", 1e-3), ("vyd", 1e-3), ("vzd", 1e-3)):
        n, err = close(g[k], ref[k], sc)
        assert n == 0, f"{k}: {n} values beyond 1e-12 (worst {err:.2e})"


@pytest.mark.parametrize("case", ["co_one_dimer_restart", "co_two_dimers_restart_sort"])
def test_restart_through_data_h5_equals_reference(golden, tmp_path, case):
    """The reference's own restart files: data.h5 (its solvent and dimer
    frames written at steps 0 and restart_at by main_amd/libmpcx_h5.so),
    timestep.dat and mt_state.dat; the restarted run's integer state equals
    the reference's restarted run (the fixture's hashes)."""
    from main_amd import h5data
    cs, p, s, tour = _inputs(case, golden)
    ra, steps, S = cs["restart_at"], cs["steps"], p["MPCtstep"]
    freqs = (ra, ra, S, S)
    slabs = tuple(float(b) / 2.0 for b in p["box"])
    stream = MtmodStream(mt=s["mt"], mti=s["mti"])
    a = CollectiveEngine(p, rng_mode=0)
    path = str(tmp_path / "data.h5")
    try:
        a.set_state(s, tour=tour)
        a.setup(do_sort=cs["sort"])
        h5data.initiate(path, p, freqs)
        h5data.write_engine(path, a, freqs, slabs)
        for _ in range(ra // S):
            a.run(1, host_draws(stream, p, 1))
            h5data.write_engine(path, a, freqs, slabs)
        o = a.observables()
        RS.write_timestep(str(tmp_path / RS.TIMESTEP), ra, int(o["nA"]), int(o["nB"]), int(o["nS"]))
        RS.write_mt_state(str(tmp_path / RS.MT_STATE), [stream])
    finally:
        a.close()
    b = CollectiveEngine(p, rng_mode=0)
    try:
        r = RS.restart_from_h5(b, str(tmp_path), s["react_box"], freqs, tour=tour)
        assert r["cstep"] == ra
        b.run((steps - ra) // S, host_draws(r["streams"][0], p, (steps - ra) // S))
        g = b.get_state()
        cells = b.get_cells()
    finally:
        b.close()
    rec = golden["cases"][case]
    got = {"flag": _h(g["flag"].astype(np.int32)), "ncn": _h(cells["ncn"].astype(np.int64)),
           "tcs": _h(cells["tcs"]), "tA": _h(cells["tA"][: p["vol"]]), "tB": _h(cells["tB"][: p["vol"]]),
           "tS": _h(cells["tS"][: p["vol"]])}
    bad = [k for k, v in got.items() if v != rec["arrays"][k]]
    assert not bad, f"{case}: differs from the reference's restarted run in {bad}"
    for k in ("cstep", "time_count", "time_penetrate"):
        assert g[k] == rec["scalars"][k], k


@pytest.mark.parametrize("case,near_mode", [("co_one_dimer_restart", None), ("co_two_dimers_restart_sort", 2),
                                            ("co_two_dimers_restart_sort", 0), ("co_six_dimers_restart", None)])
def test_restart_evaluation_matches_oracle(golden, case, near_mode):
    """Forces right after the restart set-up, from a mid-run state with no
    forces: dimer forces, every pair force of every list, the potential and
    the species (A inside a catalytic cutoff turns B)."""
    cs, p, s, tour = _inputs(case, golden)
    ref = dict(s)
    O.co_run(ref, 0, do_setup=2 if cs["sort"] else 1, tour=tour)
    O.co_run(ref, cs["restart_at"], tour=tour)
    mid = {k: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v) for k, v in ref.items()}
    for k in ("fxd", "fyd", "fzd", "fpx", "fpy", "fpz", "pot"):
        ref[k][...] = 0.0
    do_sort = cs["restart_at"] % int(p["sort_step"]) == 0
    O.co_run(ref, 0, do_setup=4 if do_sort else 3, tour=tour)
    eng = CollectiveEngine(p, rng_mode=0, near_mode=near_mode)
    try:
        for k in ("fxd", "fyd", "fzd"):
            mid[k] = np.zeros_like(mid[k])
        eng.set_state(mid, tour=tour)
        eng.setup(do_sort=do_sort, restart=True)
        g = eng.get_state()
        lists = eng.get_lists()
        tp = eng.tpotential()
    finally:
        eng.close()
    assert np.array_equal(g["flag"], ref["flag"])
    assert np.array_equal(lists["nbnum"], ref["nbnum"])
    nd = p["numdims"]
    for k in range(nd):
        for i in range(2):
            n = int(ref["nbnum"][k, i])
            assert np.array_equal(lists["nb"][k, i, :n], ref["nb"][k, i, :n])
            for f in ("fpx", "fpy", "fpz", "pot"):
                bad, err = close(lists[f][k, i, :n], ref[f][k, i, :n], 1.0)
                assert bad == 0, f"{f}({k},{i}): worst {err:.2e}"
    for f in ("fxd", "fyd", "fzd"):
        bad, err = close(g[f], ref[f], 1.0)
        assert bad == 0, f"{f}: worst {err:.2e}"
    assert abs(tp - ref["tpotential"]) <= 1e-12 * max(abs(ref["tpotential"]), 1.0)
    assert any(abs(float(v)) > 1e-6 for v in np.ravel(ref["fxd"]))   # the case exercises nonzero forces


def test_frame_round_trip_and_checks(tmp_path):
    from bench import headline_params
    from main_amd import systems as S
    p = headline_params(40)
    s = S.collective_initial_state(p, seed=5)
    tour = cell_tour(40, 40, 40)
    a = CollectiveEngine(p, rng_mode=1, seed=5)
    b = CollectiveEngine(p, rng_mode=1, seed=5)
    try:
        a.set_state(s, tour=tour)
        a.setup(do_sort=True)
        a.run(3)
        f = str(tmp_path / "frame.bin")
        a.save_frame(f)
        ga = a.get_state()
        assert b.load_frame(f, s["react_box"], tour) == ga["cstep"]
        gb = b.get_state()
        for k in ("x", "y", "z", "vx", "vy", "vz", "flag", "xd", "yd", "zd", "vxd", "vyd", "vzd"):
            assert np.array_equal(np.asarray(ga[k]).view(np.uint8), np.asarray(gb[k]).view(np.uint8)), k
        raw = bytearray(open(f, "rb").read())
        raw[64 + 8 * 1000] ^= 1
        open(f + ".bad", "wb").write(bytes(raw))
        with pytest.raises(MpcxError, match="checksum"):
            b.load_frame(f + ".bad", s["react_box"], tour)
    finally:
        a.close()
        b.close()
    q = headline_params(30)
    c = CollectiveEngine(q, rng_mode=1, seed=5)
    try:
        with pytest.raises(MpcxError, match="another system"):
            c.load_frame(f, np.zeros(900, np.int32), cell_tour(30, 30, 30))
    finally:
        c.close()
