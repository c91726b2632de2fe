"""data.h5 of collective_code through main_amd/libmpcx_h5.so (HDF5 C API).

    initiate(path, p, freqs)              output_mod.f95:96-264   output_hdf5_initiate
    write(path, cstep, p, state, ...)     output_mod.f95:324-898  output_hdf5_write
    write_engine(path, eng, freqs, ...)   the same frames from a CollectiveEngine's state
    read_restart(path, cstep, p, freqs)   restart_mod.f95:359-402 (the data.h5 frames)

freqs = (Freq1, Freq2, Freq3, Freq4): the output periods (in MD steps) of the
solvent, concentration, system and dimer groups (input.txt's last four lines).
The concentration slabs default to the reference's widths 0.5 a0
(main_pattern.f95:74-76) and its slab half-width cutoffC = Rc * cutoff
(:120).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpcx_h5.so")
_lib = None
I32, I64, F64 = C.c_int32, C.c_int64, C.c_double


class H5Error(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise H5Error(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        _lib = C.CDLL(LIB_PATH)
        _lib.mpcx_h5_last_error.restype = C.c_char_p
    return _lib


def _check(rc, what):
    if rc != 0:
        raise H5Error(f"{what}: {lib().mpcx_h5_last_error().decode(errors='replace')}")


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def initiate(path, p, freqs, widths=(0.5, 0.5, 0.5)):
    box = np.ascontiguousarray(p["box"], np.float64)
    _check(lib().mpcx_h5_initiate(str(path).encode(), *[I32(int(f)) for f in freqs], _p(box), F64(p["T"]),
                                  F64(p["MDt"]), F64(p["MPCt"]), I32(int(p["NT"])), I32(int(p["numdims"])),
                                  *[F64(w) for w in widths]), "mpcx_h5_initiate")


def write(path, cstep, p, state, freqs, slabs, widths=(0.5, 0.5, 0.5), tpotential=0.0):
    """state: x y z vx vy vz (NT), flag, and the (numdims, 2) dimer arrays
    xd yd zd vxd vyd vzd; slabs = (x_conc_slab, y_conc_slab, z_conc_slab)."""
    f64 = lambda k: np.ascontiguousarray(state[k], np.float64)
    dims = {k: np.asfortranarray(state[k], np.float64) for k in ("xd", "yd", "zd", "vxd", "vyd", "vzd")}
    sol = [f64(k) for k in ("x", "y", "z")]
    vel = [f64(k) for k in ("vx", "vy", "vz")]
    flag = np.ascontiguousarray(state["flag"], np.int32)
    box = np.ascontiguousarray(p["box"], np.float64)
    _check(lib().mpcx_h5_write(str(path).encode(), I32(int(cstep)), I64(int(p["vol"])), *[I32(int(f)) for f in freqs],
                               I32(int(p["NT"])), I32(int(p["numdims"])),
                               *[F64(p[k]) for k in ("Ma", "Mb", "Ms", "Mc", "Mn")], *[F64(s) for s in slabs],
                               _p(box), *[F64(w) for w in widths], *[_p(a) for a in sol], _p(flag),
                               *[_p(a) for a in vel], F64(p["Rc"] * p["cutoff"]), F64(p["Dcn"]),
                               *[_p(dims[k]) for k in ("xd", "yd", "zd", "vxd", "vyd", "vzd")], F64(tpotential)),
           "mpcx_h5_write")


def write_engine(path, eng, freqs, slabs, widths=(0.5, 0.5, 0.5)):
    """The frames due at the engine's current step, from its device state."""
    s = eng.get_state()
    write(path, s["cstep"], eng.p, s, freqs, slabs, widths, eng.tpotential())
    return s["cstep"]


def read_restart(path, cstep, p, freqs):
    """The solvent of frame cstep/Freq1 and the dimers of frame cstep/Freq4."""
    NT, nd = int(p["NT"]), int(p["numdims"])
    out = {k: np.empty(NT) for k in ("x", "y", "z", "vx", "vy", "vz")}
    out["flag"] = np.empty(NT, np.int32)
    for k in ("xd", "yd", "zd", "vxd", "vyd", "vzd"):
        out[k] = np.empty((nd, 2), order="F")
    _check(lib().mpcx_h5_read_restart(str(path).encode(), I32(int(cstep)), I32(int(freqs[0])), I32(int(freqs[3])),
                                      I32(NT), I32(nd), *[_p(out[k]) for k in ("x", "y", "z")], _p(out["flag"]),
                                      *[_p(out[k]) for k in ("vx", "vy", "vz")],
                                      *[_p(out[k]) for k in ("xd", "yd", "zd", "vxd", "vyd", "vzd")]),
           "mpcx_h5_read_restart")
    return out
