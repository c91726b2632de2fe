"""Restart files and text records against the reference's own I/O.

The engine library's codecs (host code, no GPU needed) are compared with:
* the reference's mtsavef/mtgetf (mtmod.f90:152-229), compiled from
  /root/reference into oracle/_ref, byte for byte;
* the Fortran runtime's list-directed and formatted I/O running the
  reference's own statements (main_pattern.f95:130-136,243-249,
  initial_mod.f95:118, restart_mod.f95:410-414, the FORMAT statements of
  Output_SelkovDimer.f90 and Main_SelkovDimer.f90), also in oracle/_ref;
* fixed strings of the gfortran layout and the record structure of the
  unformatted file, which need no reference build.
The oracle's restart evaluation (restart_mod.f95:425-495) is pinned by the
restart cases of tests/golden (reference hashes); the check below shows those
cases depend on it.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from main_amd import restart as RS
from main_amd import selkov_files as SF
from main_amd._lib import MtmodStream
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FCHECK = os.path.join(ROOT, "tests", "fortran", "_build", "restart_files_check")
REF_CO = O.ref_lib("collective")
REF_SK = O.ref_lib("selkov")
need_ref_co = pytest.mark.skipif(REF_CO is None, reason="oracle/_ref not built (needs /root/reference)")
need_ref_sk = pytest.mark.skipif(REF_SK is None, reason="oracle/_ref not built (needs /root/reference)")


def _path(p):
    b = os.fsencode(str(p))
    return b, C.c_int32(len(b))


def _awkward_doubles(n=60, seed=1):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=n) * 10.0 ** rng.integers(-120, 120, n)
    special = [0.0, -0.0, 1.0, -1.0, 0.1, 0.09999999999999999, 1e16, 9.999999999999999e16, 1e17, 5e-324,
               1.7976931348623157e308, -12.345678901234567, 123456.789, 2.0 / 3.0]
    return np.concatenate([special, v])


# ------------------------------------------------------------ mt_state.dat
def _streams(T):
    out = []
    for t in range(T):
        s = MtmodStream(seed=4357 + 17 * t)
        s.draws(100 + 300 * t)     # different mti per stream
        out.append(s)
    return out


def test_mt_state_record_structure(tmp_path):
    T = 3
    st = _streams(T)
    f = tmp_path / "mt_state.dat"
    RS.write_mt_state(str(f), st)
    b = f.read_bytes()
    assert len(b) == (4 + 4 * T + 4) + (4 + 2496 * T + 4)
    w = np.frombuffer(b, np.uint32)
    assert w[0] == w[1 + T] == 4 * T
    assert w[2 + T] == w[-1] == 2496 * T
    assert list(w[1:1 + T]) == [s.mti.value for s in st]
    mt = w[3 + T:-1].reshape(T, 624)       # mt(1:624, t) column after column
    for t in range(T):
        assert np.array_equal(mt[t], st[t].mt)
    back = RS.read_mt_state(str(f))
    for a, s in zip(back, st):          # the restored streams continue identically
        assert a.mti.value == s.mti.value
        assert np.array_equal(a.draws(1000), s.draws(1000))


@need_ref_co
@pytest.mark.parametrize("T", [1, 4])
def test_mt_state_matches_reference_mtsavef(tmp_path, T):
    st = _streams(T)
    ours, ref = tmp_path / "ours.dat", tmp_path / "ref.dat"
    RS.write_mt_state(str(ours), st)
    mt = np.asfortranarray(np.stack([s.mt for s in st], axis=1).view(np.int32))
    mti = np.array([s.mti.value for s in st], np.int32)
    REF_CO.ref_mtsavef(*_path(ref), C.c_int32(T), O._ptr(mt), O._ptr(mti))
    assert ours.read_bytes() == ref.read_bytes()
    # the reference's mtgetf reads ours, ours reads the reference's
    mt2 = np.zeros((624, T), np.int32, order="F")
    mti2 = np.zeros(T, np.int32)
    REF_CO.ref_mtgetf(*_path(ours), C.c_int32(T), O._ptr(mt2), O._ptr(mti2))
    assert np.array_equal(mt2, mt) and np.array_equal(mti2, mti)
    back = RS.read_mt_state(str(ref))
    assert [s.mti.value for s in back] == list(mti)


def test_mt_state_rejects_damaged_files(tmp_path):
    f = tmp_path / "mt_state.dat"
    RS.write_mt_state(str(f), _streams(2))
    b = bytearray(f.read_bytes())
    (tmp_path / "short.dat").write_bytes(bytes(b[:-8]))
    b2 = bytearray(b)
    b2[4 + 8 + 4] ^= 1                      # second record marker
    (tmp_path / "marker.dat").write_bytes(bytes(b2))
    for name in ("short.dat", "marker.dat"):
        with pytest.raises(Exception):
            RS.read_mt_state(str(tmp_path / name))


# ------------------------------------------------------------ timestep.dat
def test_timestep_gfortran_layout(tmp_path):
    f = tmp_path / "timestep.dat"
    RS.write_timestep(str(f), 2000, 123456, 7, 0)
    assert f.read_text() == "        2000\n" + "%21d\n" % 123456 + "%21d\n" % 7 + "%21d\n" % 0
    assert RS.read_timestep(str(f)) == (2000, 123456, 7, 0)


def test_timestep_reads_other_list_directed_forms(tmp_path):
    f = tmp_path / "timestep.dat"
    f.write_text(" 40 , ignored rest\n\n  1000000\n2*77\n 5 / \n")   # blank record, repeat count, slash
    # each READ starts a fresh record: 40 | 1000000 | 77 (first of 2*77) | 5
    assert RS.read_timestep(str(f)) == (40, 1000000, 77, 5)
    f.write_text("12\n3\n")
    with pytest.raises(Exception):
        RS.read_timestep(str(f))


@need_ref_co
def test_timestep_against_fortran_list_directed(tmp_path):
    ours = tmp_path / "ours.dat"
    RS.write_timestep(str(ours), 123450, 2 ** 40 + 3, 17, 9)
    cs = np.zeros(1, np.int32)
    tot = np.zeros(3, np.int64)
    REF_CO.ref_timestep_read(*_path(ours), O._ptr(cs), O._ptr(tot))
    assert int(cs[0]) == 123450 and list(tot) == [2 ** 40 + 3, 17, 9]
    ref = tmp_path / "ref.dat"
    REF_CO.ref_timestep_write(*_path(ref), C.c_int32(-31), O._ptr(np.array([5, 0, 2 ** 33], np.int64)))
    assert RS.read_timestep(str(ref)) == (-31, 5, 0, 2 ** 33)


# ------------------------------------------------------- motor_positions.dat
def test_motor_positions_gfortran_layout_and_round_trip(tmp_path):
    f = tmp_path / "motor_positions.dat"
    RS.write_motor_positions(str(f), np.array([[1.0, 0.001, -12.345678901234567]]))
    assert f.read_text() == ("   1.0000000000000000     " + "   1.0000000000000000E-003"
                             + "  -12.345678901234567     " + "\n")
    v = _awkward_doubles()
    nd = v.size // 3
    cm = v[: 3 * nd].reshape(nd, 3)
    RS.write_motor_positions(str(f), cm)
    back = RS.read_motor_positions(str(f), nd)
    assert np.array_equal(back.view(np.uint64), np.asfortranarray(cm).view(np.uint64))


def test_motor_positions_reads_other_forms(tmp_path):
    f = tmp_path / "motor_positions.dat"
    f.write_text(" 1.5D0, 2*-3.25\n 4.E+1 .5 6.0-001\n")
    back = RS.read_motor_positions(str(f), 2)
    assert np.array_equal(back, [[1.5, -3.25, -3.25], [40.0, 0.5, 0.6]])


@need_ref_co
def test_motor_positions_against_fortran_list_directed(tmp_path):
    v = _awkward_doubles(seed=7)
    nd = v.size // 3
    cm = np.asfortranarray(v[: 3 * nd].reshape(nd, 3))
    ours = tmp_path / "ours.dat"
    RS.write_motor_positions(str(ours), cm)
    back = np.zeros((nd, 3), order="F")
    REF_CO.ref_motor_positions_read(*_path(ours), C.c_int32(nd), O._ptr(back))   # restart_mod.f95:410-414
    assert np.array_equal(back.view(np.uint64), cm.view(np.uint64))
    ref = tmp_path / "ref.dat"
    REF_CO.ref_motor_positions_write(*_path(ref), C.c_int32(nd), O._ptr(cm))   # initial_mod.f95:118
    assert np.array_equal(RS.read_motor_positions(str(ref), nd).view(np.uint64), cm.view(np.uint64))


def test_motor_cm_is_sphere_midpoint():
    s = {"xd": np.array([[1.0, 3.0]]), "yd": np.array([[2.0, 2.0]]), "zd": np.array([[0.0, -4.0]])}
    assert np.array_equal(RS.motor_cm(s), [[2.0, 2.0, -2.0]])


# -------------------------------------------------------- fixed-form records
def test_fixed_form_known_records():
    assert SF.format_record("E23.16", [1.0]) == "  0.1000000000000000E+01"[1:]
    assert SF.format_record("S23.16", [-1.0]) == "-1.0000000000000000E+00"
    assert SF.format_record("E23.16", [0.0]) == " 0.0000000000000000E+00"
    assert SF.format_record("S23.16", [1e100]) == " 1.0000000000000000+100"
    assert SF.format_record("I8 X1 S23.16 X1", [42, 2.5]) == "      42  2.5000000000000000E+00"
    assert SF.format_record("I9", [7]) == "        7"
    assert SF.format_record("I2", [123]) == "**"


@need_ref_sk
def test_fixed_form_records_match_fortran_format(tmp_path):
    rng = np.random.default_rng(11)
    f = tmp_path / "rec.dat"
    for t in range(100):
        v = rng.normal(size=10) * 10.0 ** rng.integers(-120, 120, 10)
        if t == 0:
            v[:] = [0.0, -0.0, 1.0, 0.1, -1.0, 1e100, 1e-100, 1e99, 123456.789, -5e-324]
        cstep = int(rng.integers(0, 10 ** 8))
        REF_SK.ref_sk_records(*_path(f), C.c_int32(cstep), O._ptr(v))
        ours = "".join(SF.format_record(spec, vals) + "\n" for spec, vals in (
            (SF.FMT15, [cstep, *v[:3]]), (SF.FMT16, [*v[:3], cstep]), (SF.FMT18, [cstep, *v[:4]]),
            (SF.FMT19, [cstep, *v]), (SF.FMT17, [cstep, *v[:8]]), (SF.FMT23, [cstep])))
        assert ours == f.read_text(), t


def test_dump_names_follow_the_reference():
    assert SF.dump_name("SolventFPos", 5) == "SolventFPos 5.dat"         # I2 of a one-digit step
    assert SF.dump_name("SolventFPos", 100) == "SolventFPos100.dat"
    assert SF.dump_name("SolventGVel", 1234567) == "SolventGVel1234567.dat"
    assert SF.restart_dump_name("SolventFPos", 99) == "SolventFPos 99.dat"  # Restart.f90 widths by pstep+1
    assert SF.restart_dump_name("SolventFPos", 1000) == "SolventFPos1000.dat"


def test_selkov_restart_files_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    NT = 500
    flag = rng.choice([0, 1, 3], NT)
    s = {k: rng.normal(size=NT) * 10.0 ** rng.integers(-110, 3, NT) for k in
         ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz")}
    s.update(flag=flag, NT=NT, dm=np.asfortranarray(rng.normal(size=(2, 12))), tpotential=1.5)
    cstep = 1230
    files = SF.dimer_records(s, cstep)
    files.update(SF.solvent_dumps(s, cstep))
    SF.write_files(str(tmp_path), files)
    SF.write_timestep(str(tmp_path), cstep, tuple(int(np.sum(flag == f)) for f in (0, 1, 3)))
    r = SF.read_restart(str(tmp_path), {})
    order = np.concatenate([np.nonzero(flag == f)[0] for f in (0, 1, 3)])   # Restart.f90 order: F, G, I
    assert r["cstep"] == cstep and r["NT"] == NT
    assert np.array_equal(r["flag"], flag[order])
    for k in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz"):
        assert np.array_equal(r[k].view(np.uint64), s[k][order].view(np.uint64)), k
    # E23.16 keeps 16 significant digits: the dimer comes back rounded to them, as in the reference
    want = np.vectorize(lambda v: float("%.15e" % v))(s["dm"][:, :9])
    assert np.array_equal(r["dm"][:, :9], want)


def test_energy_momentum_sums_in_particle_order():
    rng = np.random.default_rng(4)
    NT = 300
    s = {k: rng.normal(size=NT) for k in ("vx", "vy", "vz")}
    s.update(flag=rng.choice([0, 1, 2, 3], NT), NT=NT, dm=np.asfortranarray(rng.normal(size=(2, 12))),
             tpotential=0.25)
    p = dict(Mf=1.0, Mg=2.0, Mi=0.5, Mc=10.0, Mn=30.0)
    got = SF.energy_momentum(s, p)
    # Output_SelkovDimer.f90:246-299 as a plain particle loop
    spx = spy = spz = ke = 0.0
    n = 0.0
    for i in range(NT):
        f = s["flag"][i]
        if f in (0, 1, 3):
            m = {0: 1.0, 1: 2.0, 3: 0.5}[f]
            px, py, pz = m * s["vx"][i], m * s["vy"][i], m * s["vz"][i]
            spx, spy, spz = spx + px, spy + py, spz + pz
            ke = ke + 0.5 * (s["vx"][i] * px + s["vy"][i] * py + s["vz"][i] * pz)
            n += 1.0
    dm = s["dm"]
    vav = [(10.0 * dm[0, 3 + a] + 30.0 * dm[1, 3 + a]) / (30.0 + 10.0) for a in range(3)]
    pd = [40.0 * v for v in vav]
    ked = 0.5 * (1 / 40.0) * (pd[0] * pd[0] + pd[1] * pd[1] + pd[2] * pd[2])
    want = [spx + pd[0], spy + pd[1], spz + pd[2], ked + ke, 0.25, ked + ke + 0.25,
            (ke / n) * float(np.float32(2) / np.float32(3)), ked * float(np.float32(2) / np.float32(3))]
    assert got == want


# ------------------------------------------------- the restart golden cases
def test_restart_cases_depend_on_the_restart_forces():
    """A set-up that kept the zero forces of a frame (a plain set-up) gives a
    different trajectory, so the reference hashes of the restart cases pin
    the restart evaluation itself."""
    from main_amd.tour import tour_coords
    from tests.golden.make_golden import COLLECTIVE_CASES, collective_inputs
    case = COLLECTIVE_CASES["co_one_dimer_restart"]
    out = []
    for mode in (3, 1):
        p, s = collective_inputs(case)
        O.co_run(s, 0, do_setup=1)
        O.co_run(s, case["restart_at"])
        for k in ("fxd", "fyd", "fzd", "fpx", "fpy", "fpz", "pot"):
            s[k][...] = 0.0
        O.co_run(s, 0, do_setup=mode)
        O.co_run(s, case["steps"] - case["restart_at"])
        out.append(s)
    assert np.max(np.abs(out[0]["vx"] - out[1]["vx"])) > 1e-6
    assert np.max(np.abs(out[0]["xd"] - out[1]["xd"])) > 1e-7


@pytest.mark.skipif(not os.path.exists(FCHECK), reason="Fortran check not built (make -C main_amd/fortran)")
def test_fortran_statements_read_and_write_the_library_files(tmp_path):
    """A flang program (tests/fortran/restart_files_check.f90) calls the codecs
    through the ISO_C_BINDING interfaces and reads/writes the same files with
    the reference's own list-directed and unformatted statements."""
    r = subprocess.run([FCHECK], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "restart files ok" in r.stdout, r.stdout + r.stderr
