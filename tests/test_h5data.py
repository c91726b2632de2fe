"""data.h5 (main_amd/libmpcx_h5.so) against the reference's layout and values.

The reference's writer (collective_code/output_mod.f95) goes through HDF5's
Fortran module, built here for no Fortran compiler with that module, so it
cannot run; the file is checked instead against what that code fixes:

* layout, read back through the HDF5 C library directly (not through the
  writer's own reader): the groups, datasets and attributes named in
  output_mod.f95:19-93 (trailing blanks trimmed), element types, the Fortran
  shapes reversed into C order, unlimited dimensions, the chunk shapes, and the
  2-D concentration fields' initial extent int(box*width) of :123-126;
* values: the system sums against tests/diag_reference.observables (pinned
  to the reference's output routines by test_diag_pins.py), the
  concentration fields against its slab counts divided by the slab volumes of
  :501, :524, :547, :570, the solvent and dimer frames equal to the arrays
  written, and the dimer unit vectors of :803-821;
* restart: restart_mod.f95:359-402 reads the frames back (the writer's reader
  against the arrays written, at frames 0 and 2).
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np
import pytest

from main_amd import systems as S
from tests import diag_reference as R

H5LIB = "/opt/conda/lib/libhdf5.so"
FREQS = (100, 100, 50, 50)
SLABS = (10.0, 8.0, 6.0)


def _h5():
    if not os.path.exists(H5LIB):
        pytest.skip("HDF5 library absent")
    from main_amd import h5data
    try:
        h5data.lib()
    except Exception as exc:
        pytest.skip(f"libmpcx_h5.so unavailable: {exc}")
    h = C.CDLL(H5LIB)
    h.H5open()
    for f in ("H5Fopen", "H5Dopen2", "H5Dget_space", "H5Gopen2", "H5Aopen", "H5Dget_type", "H5Dget_create_plist",
              "H5Aget_space", "H5Aget_type"):
        getattr(h, f).restype = C.c_int64
    return h5data, h


class Reader:
    """Datasets and attributes read with the HDF5 C library."""

    def __init__(self, h, path):
        self.h = h
        self.f = h.H5Fopen(str(path).encode(), C.c_uint(0), C.c_int64(0))
        assert self.f >= 0
        self.dbl = C.c_int64.in_dll(h, "H5T_NATIVE_DOUBLE_g").value
        self.int = C.c_int64.in_dll(h, "H5T_NATIVE_INT_g").value

    def dims(self, name):
        d = self.h.H5Dopen2(C.c_int64(self.f), name.encode(), C.c_int64(0))
        assert d >= 0, name
        sp = self.h.H5Dget_space(C.c_int64(d))
        n = self.h.H5Sget_simple_extent_ndims(C.c_int64(sp))
        dims, maxd = (C.c_uint64 * n)(), (C.c_uint64 * n)()
        self.h.H5Sget_simple_extent_dims(C.c_int64(sp), dims, maxd)
        pl = self.h.H5Dget_create_plist(C.c_int64(d))
        ch = (C.c_uint64 * n)()
        self.h.H5Pget_chunk(C.c_int64(pl), n, ch)
        t = self.h.H5Dget_type(C.c_int64(d))
        is_int = self.h.H5Tequal(C.c_int64(t), C.c_int64(self.int)) > 0
        self.h.H5Pclose(C.c_int64(pl))
        self.h.H5Tclose(C.c_int64(t))
        self.h.H5Sclose(C.c_int64(sp))
        self.h.H5Dclose(C.c_int64(d))
        return [list(dims), list(maxd), list(ch), is_int]

    def read(self, name, is_int=False):
        dims = self.dims(name)[0]
        out = np.empty(dims, np.int32 if is_int else np.float64)
        d = self.h.H5Dopen2(C.c_int64(self.f), name.encode(), C.c_int64(0))
        rc = self.h.H5Dread(C.c_int64(d), C.c_int64(self.int if is_int else self.dbl), C.c_int64(0), C.c_int64(0),
                            C.c_int64(0), out.ctypes.data_as(C.c_void_p))
        self.h.H5Dclose(C.c_int64(d))
        assert rc >= 0, name
        return out

    def attr(self, obj, name, is_int=False):
        o = self.h.H5Gopen2(C.c_int64(self.f), obj.encode(), C.c_int64(0)) if obj != "/" else self.f
        a = self.h.H5Aopen(C.c_int64(o), name.encode(), C.c_int64(0))
        assert a >= 0, (obj, name)
        sp = self.h.H5Aget_space(C.c_int64(a))
        n = max(self.h.H5Sget_simple_extent_npoints(C.c_int64(sp)), 1)
        out = np.empty(n, np.int32 if is_int else np.float64)
        t = self.h.H5Aget_type(C.c_int64(a))
        stored_int = self.h.H5Tequal(C.c_int64(t), C.c_int64(self.int)) > 0
        self.h.H5Aread(C.c_int64(a), C.c_int64(self.int if is_int else self.dbl), out.ctypes.data_as(C.c_void_p))
        self.h.H5Tclose(C.c_int64(t))
        self.h.H5Sclose(C.c_int64(sp))
        self.h.H5Aclose(C.c_int64(a))
        if obj != "/":
            self.h.H5Gclose(C.c_int64(o))
        return out, stored_int

    def close(self):
        self.h.H5Fclose(C.c_int64(self.f))


def _system(nd=2, seed=4):
    p = S.collective_params(Lx=24.0, Ly=20.0, Lz=16.0, numdims=nd, rhoa=4.0, rhob=1.0)
    s = S.collective_initial_state(p, seed=seed)
    rng = np.random.default_rng(seed)
    s["flag"][:] = rng.choice(np.array([0, 1, 2], np.int32), p["NT"], p=[0.6, 0.3, 0.1])
    return p, s


def _frames(h5data, path, p, s, steps):
    """Frames at the given steps, the state changed between them."""
    h5data.initiate(path, p, FREQS)
    states = []
    rng = np.random.default_rng(9)
    for k, cstep in enumerate(steps):
        st = {key: np.array(v, copy=True) for key, v in s.items() if isinstance(v, np.ndarray)}
        if k:
            for key in ("vx", "vy", "vz", "vxd", "vyd", "vzd"):
                st[key] = st[key] + rng.normal(0, 0.01, st[key].shape)
        tp = 1.5 + k
        h5data.write(path, cstep, p, st, FREQS, SLABS, tpotential=tp)
        states.append((cstep, st, tp))
    return states


def test_layout_is_the_reference_layout(tmp_path):
    h5data, h = _h5()
    p, s = _system()
    path = tmp_path / "data.h5"
    _frames(h5data, path, p, s, [0])
    r = Reader(h, path)
    try:
        NT, nd = p["NT"], p["numdims"]
        box, _ = r.attr("/", "simulation_box_length")
        assert list(box) == list(p["box"])
        for name, val in (("solvent_temperature", p["T"]), ("MD_time", p["MDt"]), ("MPC_time", p["MPCt"]),
                          ("total_#_particles", NT), ("number_of_dimers", nd)):
            v, stored_int = r.attr("/", name)
            assert v[0] == val and not stored_int, name   # created double, also for the counts
        for grp, f in (("Solvent", FREQS[0]), ("Concentration", FREQS[1]), ("System", FREQS[2]),
                       ("Dimer", FREQS[3])):
            v, stored_int = r.attr(grp, "io_frequency", is_int=True)
            assert v[0] == f and stored_int, grp
        U = 2 ** 64 - 1   # H5S_UNLIMITED
        for k in ("position_x", "position_y", "position_z", "flag", "velocity_x", "velocity_y", "velocity_z"):
            dims, maxd, ch, is_int = r.dims("Solvent/" + k)
            assert dims == [1, NT] and maxd == [U, U] and ch == [1, 100000] and is_int == (k == "flag"), k
        for k in ("pxsol", "pysol", "pzsol", "total_energy", "solvent_temperature", "dimer_temperature"):
            assert r.dims("System/" + k)[:3] == [[1], [U], [75000]], k
        nx, ny, nz = (int(b / 0.5) for b in p["box"])
        qx, qy, qz = (int(b * 0.5) for b in p["box"])
        assert r.dims("Concentration/total_A")[:3] == [[1], [U], [75000]]
        assert r.dims("Concentration/1D_concentration_A")[:3] == [[1, nx], [U, U], [800, 120]]
        assert r.dims("Concentration/2D_concentration_A_xy")[:3] == [[1, qy, qx], [U, U, U], [7, 120, 120]]
        assert r.dims("Concentration/2D_concentration_B_xz")[0] == [1, qz, qx]
        assert r.dims("Concentration/2D_concentration_A_yz")[0] == [1, qz, qy]
        for k in ("position_x", "velocity_z"):
            assert r.dims("Dimer/" + k)[:3] == [[1, 2, nd], [U, U, U], [math.floor(100000.0 / (2 * nd)), 2, nd]]
        assert r.dims("Dimer/nucvec")[:3] == [[1, 3, nd], [U, U, U], [math.floor(100000.0 / (3 * nd)), 3, nd]]
    finally:
        r.close()


def test_frames_hold_the_reference_values(tmp_path):
    h5data, h = _h5()
    p, s = _system()
    path = tmp_path / "data.h5"
    states = _frames(h5data, path, p, s, [0, 50, 100, 150])
    r = Reader(h, path)
    try:
        nd, NT = p["numdims"], p["NT"]
        sysd = {k: r.read("System/" + k) for k in ("pxsol", "pysol", "pzsol", "total_energy", "solvent_temperature",
                                                  "dimer_temperature")}
        assert sysd["pxsol"].shape == (4,)
        for fr, (cstep, st, tp) in enumerate(states):
            q = dict(st, xd=np.stack([np.r_[st["xd"][:, 0], st["xd"][:, 1]], np.r_[st["yd"][:, 0], st["yd"][:, 1]],
                                      np.r_[st["zd"][:, 0], st["zd"][:, 1]]], axis=1),
                     vd=np.stack([np.r_[st["vxd"][:, 0], st["vxd"][:, 1]], np.r_[st["vyd"][:, 0], st["vyd"][:, 1]],
                                  np.r_[st["vzd"][:, 0], st["vzd"][:, 1]]], axis=1))
            o = R.observables(q, {0: p["Ma"], 1: p["Mb"], 2: p["Ms"]}, {0: 0, 1: 1, 2: 2}, (p["Mc"], p["Mn"]), nd,
                              p["box"], False, tp)
            assert [sysd[k][fr] for k in ("pxsol", "pysol", "pzsol")] == list(o["ptot"])
            assert sysd["total_energy"][fr] == o["E_tot"]
            assert sysd["solvent_temperature"][fr] == o["T_sol"] and sysd["dimer_temperature"][fr] == o["T_dim"]
        # solvent frames 0 and 1 (steps 0 and 100) are the arrays written
        vx = r.read("Solvent/velocity_x")
        fl = r.read("Solvent/flag", is_int=True)
        assert vx.shape == (2, NT) and np.array_equal(vx[1], states[2][1]["vx"]) and np.array_equal(vx[0], s["vx"])
        assert np.array_equal(fl[1], s["flag"])
        # dimers and unit vectors of every frame
        dx, nuc = r.read("Dimer/velocity_y"), r.read("Dimer/nucvec")
        for fr, (cstep, st, _) in enumerate(states):
            assert np.array_equal(dx[fr], st["vyd"].T)
            for j in range(nd):
                d = [st[k][j, 0] - st[k][j, 1] for k in ("xd", "yd", "zd")]
                for a in range(2):
                    L = p["box"][a]
                    d[a] = d[a] - L if d[a] > 0.5 * L else (d[a] + L if d[a] < -0.5 * L else d[a])
                assert [nuc[fr, a, j] for a in range(3)] == [v / p["Dcn"] for v in d]
        # concentration fields at step 100 (frame 1, full extent) and totals
        half = p["Rc"] * p["cutoff"]
        sl = R.slabs(states[2][1], p["box"], *SLABS, half, 0.5)
        a1 = r.read("Concentration/1D_concentration_A")
        bxz = r.read("Concentration/2D_concentration_B_xz")
        ayz = r.read("Concentration/2D_concentration_A_yz")
        assert np.array_equal(a1[1], sl["A1d"] / ((0.5 * p["box"][1]) * (2.0 * half)))
        assert np.array_equal(bxz[1], (sl["Bxz"] / ((0.5 * 0.5) * (2.0 * half))).T)
        assert np.array_equal(ayz[1], (sl["Ayz"] / ((0.5 * 0.5) * (2.0 * half))).T)
        ta = r.read("Concentration/total_A")
        assert ta[1] == float(np.sum(s["flag"] == 0))
        assert r.read("Concentration/concentration_B")[1] == float(np.sum(s["flag"] == 1)) / float(p["vol"])
        # frame 0 of a 2-D field: the leading int(Lx/2)*int(Ly/2) values in memory order
        axy = r.read("Concentration/2D_concentration_A_xy")
        sl0 = R.slabs(states[0][1], p["box"], *SLABS, half, 0.5)
        full0 = (sl0["Axy"] / ((0.5 * 0.5) * (2.0 * half))).ravel(order="F")
        qx, qy = int(p["box"][0] * 0.5), int(p["box"][1] * 0.5)
        assert np.array_equal(axy[0, :qy, :qx].ravel(), full0[: qx * qy])
        assert np.all(axy[0, qy:, :] == 0.0) and np.all(axy[0, :, qx:] == 0.0)
    finally:
        r.close()


def test_restart_reads_the_frames_back(tmp_path):
    h5data, _ = _h5()
    p, s = _system(nd=1, seed=7)
    path = tmp_path / "data.h5"
    states = _frames(h5data, path, p, s, [0, 50, 100, 150, 200])
    for cstep in (0, 200):
        got = h5data.read_restart(path, cstep, p, FREQS)
        st = {c: v for c, v, _ in states}[cstep]
        for k in ("x", "y", "z", "vx", "vy", "vz", "flag", "xd", "yd", "zd", "vxd", "vyd", "vzd"):
            assert np.array_equal(got[k], st[k]), (cstep, k)
    with pytest.raises(h5data.H5Error, match="not in the file"):
        h5data.read_restart(path, 300, p, FREQS)
