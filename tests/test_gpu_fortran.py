"""The Fortran boundary: collective_code's module procedures, replaced by the
drop-in modules of main_amd/fortran (same module names, procedure names and
dummy arguments) and compiled with flang, against the oracle.

tests/fortran/dropin_check.f90 runs sort_hilbert, neighbour_list (every
sphere), inout_particle_list, distribute, mpc_rotation and reaction_simple in
the order of main_pattern.f95's set-up and first collision, drawing its
Mersenne-Twister numbers from the caller's mt/mti like the reference.  The
oracle performs the same calls; integer results and the RNG state must be
identical, velocities within 1e-12 relative (device cos/sin of the axes).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

from main_amd.tour import cell_tour, write_hilbert_txt
from oracle import oracle as O
from tests.parity_cases import build_collective, close, small_collective_params

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "fortran", "_build", "dropin_check")
MAX_NB = 4000
P = O._ptr


def _oracle(p, s, tour):
    lib = O.oracle_lib()
    NT, nd, vol, mc, maxs = p["NT"], p["numdims"], p["vol"], p["mpcnumcell"], O.MAXS_CO
    r = {k: s[k].copy() for k in ("x", "y", "z", "vx", "vy", "vz", "flag")}
    lib.or_co_sort_hilbert(C.c_int64(vol), C.c_int32(NT), P(p["box"]), P(tour), P(r["x"]), P(r["y"]), P(r["z"]),
                           P(r["flag"]), P(r["vx"]), P(r["vy"]), P(r["vz"]))
    nb = np.zeros((nd, 2, MAX_NB), np.int64, order="F")
    nbnum = np.zeros((nd, 2), np.int64, order="F")
    for k in range(nd):
        for i, R in enumerate((p["Rc"], p["Rn"])):
            n = C.c_int64(0)
            v = nb[k, i, :]
            lib.or_co_neighbour_list(C.c_int32(NT), P(p["box"]), C.c_double(p["cutoff"] * R), C.c_double(p["skinrad"]),
                                     P(r["x"]), P(r["y"]), P(r["z"]), C.c_double(s["xd"][k, i]),
                                     C.c_double(s["yd"][k, i]), C.c_double(s["zd"][k, i]), C.byref(n),
                                     C.c_void_p(v.ctypes.data), C.c_int64(v.strides[0] // 8))
            nbnum[k, i] = n.value
    on, inn = C.c_int64(0), C.c_int64(0)
    ol, il = np.zeros(NT, np.int64), np.zeros(NT, np.int64)
    xd, yd, zd = (np.asfortranarray(s[k]) for k in ("xd", "yd", "zd"))
    lib.or_co_inout(C.c_int32(NT), P(p["box"]), C.c_double(p["cutoff"] * p["Rc"]), C.c_double(p["cutoff"] * p["Rn"]),
                    C.c_double(p["skinrad"]), P(r["x"]), P(r["y"]), P(r["z"]), P(xd), P(yd), P(zd), C.c_int32(nd),
                    C.byref(on), P(ol), C.byref(inn), P(il))
    t = {k: np.zeros(vol, np.int64) for k in ("tA", "tB", "tS")}
    sl = {k: np.zeros((vol, maxs), np.int64, order="F") for k in ("sA", "sB", "sS")}
    assert lib.or_co_distribute_flat(C.c_int32(NT), C.c_int32(nd), P(r["flag"]), P(r["x"]), P(r["y"]), P(r["z"]),
                                     P(xd), P(yd), P(zd), P(p["box"]), C.c_int64(vol), P(t["tA"]), P(t["tB"]),
                                     P(t["tS"]), P(sl["sA"]), P(sl["sB"]), P(sl["sS"]), C.c_int64(maxs),
                                     C.c_double((p["cutoff"] * p["Rc"]) ** 2), C.c_double((p["cutoff"] * p["Rn"]) ** 2),
                                     C.c_int64(on.value), P(ol), C.c_int64(inn.value), P(il)) == 0
    cells = {k: np.zeros(mc) for k in ("vxcm", "vycm", "vzcm", "tcs", "uvx", "uvy", "uvz")}
    ncn = np.zeros(NT, np.int64)
    st = O.MtmodState()
    C.memmove(st.mt, s["mt"].ctypes.data, 624 * 4)
    st.mti = s["mti"]
    assert lib.or_co_mpc_rotation(C.c_int32(NT), C.c_double(p["rotang"]), C.c_int64(mc), C.c_double(p["a0"]),
                                  P(p["box"]), P(r["x"]), P(r["y"]), P(r["z"]), P(r["vx"]), P(r["vy"]), P(r["vz"]),
                                  P(cells["vxcm"]), P(cells["vycm"]), P(cells["vzcm"]), P(cells["tcs"]), P(ncn),
                                  P(cells["uvx"]), P(cells["uvy"]), P(cells["uvz"]), C.byref(st)) == 0
    lib.or_co_reaction_simple(C.c_double(p["MPCt"]), C.c_int64(p["rxnnumcell"]), C.c_double(p["k1"]), P(r["flag"]),
                              P(t["tB"]), P(sl["sB"]), C.byref(st))
    r.update(nbnum=nbnum, nb=nb, onum=on.value, inum=inn.value, ol=ol, il=il, ncn=ncn, **t, **sl, **cells)
    r["mt"] = np.frombuffer(bytes(st.mt), dtype=np.int32).copy()
    r["mti"] = int(st.mti)
    return r


def _read_out(path, p):
    NT, nd, vol, mc, maxs = p["NT"], p["numdims"], p["vol"], p["mpcnumcell"], O.MAXS_CO
    b = open(path, "rb").read()
    o = {}
    pos = 0

    def take(name, dtype, shape):
        nonlocal pos
        n = int(np.prod(shape))
        a = np.frombuffer(b, dtype=dtype, count=n, offset=pos)
        pos += n * np.dtype(dtype).itemsize
        o[name] = a.reshape(shape, order="F")

    for k in ("x", "y", "z", "vx", "vy", "vz"):
        take(k, np.float64, (NT,))
    take("flag", np.int32, (NT,))
    take("nbnum", np.int64, (nd, 2))
    take("nb", np.int64, (nd, 2, MAX_NB))
    take("onum", np.int64, (1,))
    take("inum", np.int64, (1,))
    take("ol", np.int64, (NT,))
    take("il", np.int64, (NT,))
    for k in ("tA", "tB", "tS"):
        take(k, np.int64, (vol,))
    for k in ("sA", "sB", "sS"):
        take(k, np.int64, (vol, maxs))
    for k in ("vxcm", "vycm", "vzcm", "tcs"):
        take(k, np.float64, (mc,))
    take("ncn", np.int64, (NT,))
    for k in ("uvx", "uvy", "uvz"):
        take(k, np.float64, (mc,))
    take("mt", np.int32, (624,))
    take("mti", np.int32, (1,))
    assert pos == len(b), "unexpected output size"
    return o


@pytest.mark.parametrize("numdims", [1, 2])
def test_fortran_dropin_modules_match_oracle(tmp_path, numdims):
    if not os.path.exists(CHECK):
        pytest.fail(f"{CHECK} is not built: run __graft_entry__.build()")
    p = small_collective_params(numdims=numdims, Lx=20.0, Ly=20.0, Lz=24.0, k1=0.4)
    s = build_collective(p, seed=13 + numdims, flags_mix=True)
    s["flag"][:] = np.random.default_rng(5).integers(0, 2, p["NT"]).astype(np.int32)
    L = [int(v) for v in p["box"]]
    write_hilbert_txt(str(tmp_path / "hilbert.txt"), *L)
    with open(tmp_path / "in.bin", "wb") as f:
        f.write(np.array([p["NT"], numdims], np.int32).tobytes())
        f.write(np.array([p["vol"], p["mpcnumcell"], p["rxnnumcell"], O.MAXS_CO], np.int64).tobytes())
        f.write(np.array(list(p["box"]) + [p["cutoff"], p["Rc"], p["Rn"], p["skinrad"], p["rotang"], p["a0"],
                                           p["MPCt"], p["k1"]], np.float64).tobytes())
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            f.write(np.ascontiguousarray(s[k], np.float64).tobytes())
        f.write(np.ascontiguousarray(s["flag"], np.int32).tobytes())
        for k in ("xd", "yd", "zd"):
            f.write(np.asfortranarray(s[k], np.float64).tobytes(order="F"))
        f.write(np.ascontiguousarray(s["mt"]).astype(np.uint32).tobytes())
        f.write(np.array([s["mti"]], np.int32).tobytes())
    print(f"[fortran] numdims={numdims} NT={p['NT']}", file=sys.stderr, flush=True)
    run = subprocess.run([CHECK], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    g = _read_out(tmp_path / "out.bin", p)
    r = _oracle(p, s, cell_tour(*L))
    for k in ("x", "y", "z", "flag", "nbnum", "nb", "ncn", "tA", "tB", "tS", "tcs", "vxcm", "vycm", "vzcm"):
        assert np.array_equal(g[k], r[k]), k
    assert int(g["onum"][0]) == r["onum"] and int(g["inum"][0]) == r["inum"]
    assert np.array_equal(g["ol"][: r["onum"]], r["ol"][: r["onum"]])
    assert np.array_equal(g["il"][: r["inum"]], r["il"][: r["inum"]])
    for k, t in (("sA", "tA"), ("sB", "tB"), ("sS", "tS")):
        m = np.arange(O.MAXS_CO)[None, :] < r[t][:, None]
        assert np.array_equal(g[k][m], r[k][m]), k
    vs = np.sqrt(p["T"] / p["Ma"])
    for k in ("vx", "vy", "vz"):
        assert close(g[k], r[k], vs)[0] == 0, k
    for k in ("uvx", "uvy", "uvz"):
        assert close(g[k], r[k], 1.0)[0] == 0, k
    assert np.array_equal(g["mt"], r["mt"]) and int(g["mti"][0]) == r["mti"]
    assert int((g["flag"] != s["flag"]).sum()) > 0   # sort and reaction changed flags


ECHECK = os.path.join(ROOT, "tests", "fortran", "_build", "engine_check")


def _run_engine_check(tmp_path, mode):
    if not os.path.exists(ECHECK):
        pytest.fail(f"{ECHECK} is not built: run __graft_entry__.build()")
    run = subprocess.run([ECHECK, mode], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert run.returncode == 0, run.stderr


def test_fortran_collective_engine_loop_matches_oracle(tmp_path):
    """The engine loop of the integration guide in Fortran (bind(C) config
    mirror, create, set_state, setup with the sort, two chunks of two
    collisions with host grnd() draws, get_state) against the oracle's run of
    the same system and stream: integer state and the caller's Mersenne-Twister
    state exact, positions and velocities within 1e-12."""
    import copy
    from main_amd.collective import make_config
    p = small_collective_params(numdims=2, Lx=20.0, Ly=20.0, Lz=24.0, k1=0.3)
    s = build_collective(p, seed=29, flags_mix=True)
    L = [int(v) for v in p["box"]]
    tour = cell_tour(*L)
    nchunk, nc = 2, 2
    cfg = make_config(p, rng_mode=0)
    rb = np.ascontiguousarray(s["react_box"], np.int32)
    with open(tmp_path / "co_in.bin", "wb") as f:
        f.write(bytes(cfg))
        f.write(np.array([rb.size, nchunk, nc], np.int64).tobytes())
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            f.write(np.ascontiguousarray(s[k], np.float64).tobytes())
        f.write(np.ascontiguousarray(s["flag"], np.int32).tobytes())
        for k in ("xd", "yd", "zd", "vxd", "vyd", "vzd", "fxd", "fyd", "fzd"):
            f.write(np.asfortranarray(s[k], np.float64).tobytes(order="F"))
        f.write(rb.tobytes())
        f.write(np.ascontiguousarray(tour, np.int64).tobytes())
        f.write(np.ascontiguousarray(s["mt"]).astype(np.uint32).tobytes())
        f.write(np.array([s["mti"]], np.int32).tobytes())
    print(f"[fortran] collective engine loop NT={p['NT']}", file=sys.stderr, flush=True)
    _run_engine_check(tmp_path, "co")
    b = open(tmp_path / "co_out.bin", "rb").read()
    NT, nd = p["NT"], p["numdims"]
    pos = 0

    def take(dtype, n):
        nonlocal pos
        a = np.frombuffer(b, dtype=dtype, count=n, offset=pos)
        pos += n * np.dtype(dtype).itemsize
        return a
    g = {k: take(np.float64, NT) for k in ("x", "y", "z", "vx", "vy", "vz")}
    g["flag"] = take(np.int32, NT)
    dim = take(np.float64, 24 * nd).reshape(12, 2 * nd)
    scal = take(np.int32, 4)
    mt, mti = take(np.int32, 624), int(take(np.int32, 1)[0])
    assert pos == len(b)
    ref = copy.deepcopy(s)
    O.co_run(ref, 0, do_setup=2, tour=tour)
    O.co_run(ref, nchunk * nc * p["MPCtstep"], tour=tour)
    assert np.array_equal(g["flag"], ref["flag"])
    vs = np.sqrt(p["T"] / p["Ma"])
    for k in ("x", "y", "z"):
        assert close(g[k], ref[k], 1.0)[0] == 0, k
    for k in ("vx", "vy", "vz"):
        assert close(g[k], ref[k], vs)[0] == 0, k
    for f, k in enumerate(("xd", "yd", "zd")):
        assert close(dim[f].reshape(2, nd).T, ref[k], 1.0)[0] == 0, k
    assert (int(scal[0]), int(scal[1]), int(scal[2])) == (ref["cstep"], ref["time_count"], ref["time_penetrate"])
    assert np.array_equal(mt.view(np.uint32), np.asarray(ref["mt"], np.uint32)) and mti == ref["mti"]


def test_fortran_selkov_engine_loop_matches_oracle(tmp_path):
    """The selkov_code loop in Fortran: bind(C) config mirror, set_state, two
    chunks of two collisions, each taking a buffer of mt95 draws and advancing
    the stream by the count the run used, against the oracle: flags and the
    stream exact, positions, velocities, forces and the dimer within 1e-12."""
    import copy
    from main_amd import systems as S
    from main_amd.selkov import make_config as sk_config
    SEED = 8154543
    p = S.selkov_params(L=20.0, rho=4.0, Rc=2.0, Rn=4.0, Dcn=7.0, MDt=0.005, MPCt=0.1, NG=3200,
                        k=(0.05, 0.01, 0.0004, 0.0004, 0.01, 0.01))
    s = S.selkov_initial_state(p, seed=9)
    s["flag"][:] = np.random.default_rng(15).choice([0, 1, 3], p["NT"]).astype(np.int64)
    nchunk, nc = 2, 2
    cfg = sk_config(p, rng_mode=0)
    space = int(cfg.space)
    with open(tmp_path / "sk_in.bin", "wb") as f:
        f.write(bytes(cfg))
        f.write(np.array([nchunk, nc], np.int64).tobytes())
        f.write(np.array([SEED], np.int32).tobytes())
        for k in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz"):
            a = np.zeros(space)
            a[: len(s[k])] = s[k]
            f.write(a.tobytes())
        fl = np.full(space, 2, np.int64)
        fl[: len(s["flag"])] = s["flag"]
        f.write(fl.tobytes())
        f.write(np.asfortranarray(s["dm"], np.float64).tobytes(order="F"))
    print(f"[fortran] selkov engine loop NT={p['NT']}", file=sys.stderr, flush=True)
    _run_engine_check(tmp_path, "sk")
    b = open(tmp_path / "sk_out.bin", "rb").read()
    pos = 0

    def take(dtype, n):
        nonlocal pos
        a = np.frombuffer(b, dtype=dtype, count=n, offset=pos)
        pos += n * np.dtype(dtype).itemsize
        return a
    g = {k: take(np.float64, space) for k in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz")}
    g["flag"] = take(np.int64, space)
    dm = take(np.float64, 24).reshape(2, 12, order="F")
    scal = take(np.int64, 4)
    take(np.float64, 1)
    take(np.int32, 625)
    assert pos == len(b)
    ref = copy.deepcopy(s)
    ref["sort_step"] = 10000
    O.sk_run(ref, 1, nchunk * nc * p["MPCtstep"], seed=SEED)
    n = int(ref["NT"])
    assert int(scal[1]) == n and int(scal[0]) == nchunk * nc * p["MPCtstep"]
    assert np.array_equal(g["flag"][:n], np.asarray(ref["flag"])[:n])
    vs = np.sqrt(p["T"] / p["Mf"])
    for k in ("x", "y", "z", "fx", "fy", "fz"):
        assert close(g[k][:n], np.asarray(ref[k])[:n], 1.0)[0] == 0, k
    for k in ("vx", "vy", "vz"):
        assert close(g[k][:n], np.asarray(ref[k])[:n], vs)[0] == 0, k
    assert close(dm[:, 0:3], np.asarray(ref["dm"])[:, 0:3], 1.0)[0] == 0
