"""Checkpoint and restart of a collective_code system on the MI355X engine.

The reference's restart protocol (main_pattern.f95:128-146 and :240-251,
restart_mod.f95) uses four files in the run directory:

* ``timestep.dat``: cstep, iTotA, iTotB, iTotS (list-directed records);
* ``mt_state.dat``: the Mersenne-Twister streams (mtsavef, unformatted);
* ``motor_positions.dat``: motorCM of every dimer, written once by initial;
* ``data.h5``: the trajectory, whose last frame restart_mod.f95:362-406 reads.

The first three are written and read here in the reference's formats (the
codecs are in the engine library).  data.h5 needs the HDF5 library, which
this build does not link: its frame is replaced by ``frame.bin``
(mpcx_co_save_frame), holding the same fields.  A restart then runs the
set-up of restart_mod.f95:416-495 on the device (``setup(restart=True)``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import MtmodStream, check, lib, ptr

TIMESTEP = "timestep.dat"
MT_STATE = "mt_state.dat"
MOTOR_POSITIONS = "motor_positions.dat"
FRAME = "frame.bin"


def _p(path: str) -> bytes:
    return os.fsencode(path)


# ----------------------------------------------------------- reference files
def write_timestep(path: str, cstep: int, n_a: int, n_b: int, n_s: int):
    """timestep.dat of main_pattern.f95:243-249."""
    check(lib().mpcx_timestep_write(_p(path), C.c_int32(cstep), C.c_int64(n_a), C.c_int64(n_b), C.c_int64(n_s)),
          "mpcx_timestep_write")


def read_timestep(path: str) -> tuple[int, int, int, int]:
    """(cstep, iTotA, iTotB, iTotS) as main_pattern.f95:130-136 reads them."""
    cs = C.c_int32(0)
    a, b, s = C.c_int64(0), C.c_int64(0), C.c_int64(0)
    check(lib().mpcx_timestep_read(_p(path), C.byref(cs), C.byref(a), C.byref(b), C.byref(s)), "mpcx_timestep_read")
    return int(cs.value), int(a.value), int(b.value), int(s.value)


def write_mt_state(path: str, streams: list[MtmodStream]):
    """mt_state.dat of mtsavef (mtmod.f90:152-177): one stream per thread."""
    T = len(streams)
    mt = np.asfortranarray(np.stack([s.mt for s in streams], axis=1).astype(np.uint32))   # (624, T)
    mti = np.array([s.mti.value for s in streams], np.int32)
    check(lib().mpcx_mt_state_write(_p(path), ptr(mt), ptr(mti), C.c_int32(T)), "mpcx_mt_state_write")


def read_mt_state(path: str) -> list[MtmodStream]:
    """The streams mtgetf (mtmod.f90:203-229) restores, thread 1 first."""
    T = C.c_int32(0)
    check(lib().mpcx_mt_state_read(_p(path), None, None, C.c_int32(0), C.byref(T)), "mpcx_mt_state_read")
    mt = np.zeros((624, T.value), np.uint32, order="F")
    mti = np.zeros(T.value, np.int32)
    check(lib().mpcx_mt_state_read(_p(path), ptr(mt), ptr(mti), T, C.byref(T)), "mpcx_mt_state_read")
    return [MtmodStream(mt=mt[:, t], mti=int(mti[t])) for t in range(T.value)]


def write_motor_positions(path: str, cm):
    """motor_positions.dat of initial_mod.f95:79,118; cm is motorCM (numdims, 3)."""
    a = np.asfortranarray(cm, dtype=np.float64)
    check(lib().mpcx_motor_positions_write(_p(path), ptr(a), C.c_int32(a.shape[0])), "mpcx_motor_positions_write")


def read_motor_positions(path: str, nd: int) -> np.ndarray:
    """motorCM (numdims, 3) as restart_mod.f95:410-414 reads it."""
    a = np.zeros((nd, 3), order="F")
    check(lib().mpcx_motor_positions_read(_p(path), ptr(a), C.c_int32(nd)), "mpcx_motor_positions_read")
    return a


def motor_cm(s: dict) -> np.ndarray:
    """motorCM of initial_mod.f95:94-96: the midpoint of the two spheres (no image)."""
    return np.asfortranarray(np.stack([0.5 * (np.asarray(s[k])[:, 0] + np.asarray(s[k])[:, 1])
                                       for k in ("xd", "yd", "zd")], axis=1))


# -------------------------------------------------------------- checkpoints
def save_checkpoint(engine, directory: str, streams: list[MtmodStream] | None = None, cm=None):
    """Write the restart files of `engine` (at a collision step) into directory.

    streams: the host Mersenne-Twister streams of an rng_mode 0 run (their
    state after the draws the run has consumed); cm: motorCM, written only if
    given (the reference writes it once, at the start of a run)."""
    os.makedirs(directory, exist_ok=True)
    engine.save_frame(os.path.join(directory, FRAME))
    o = engine.observables()
    cstep = engine.scalars()["cstep"]
    write_timestep(os.path.join(directory, TIMESTEP), cstep, int(o["nA"]), int(o["nB"]), int(o["nS"]))
    if streams:
        write_mt_state(os.path.join(directory, MT_STATE), streams)
    if cm is not None:
        write_motor_positions(os.path.join(directory, MOTOR_POSITIONS), cm)


def restart(engine, directory: str, react_box, tour=None) -> dict:
    """Restore `engine` from directory as the reference restarts: the frame,
    then the set-up of restart_mod.f95:416-495 (sort when cstep is a multiple
    of sort_step).  Returns cstep, the restored streams (empty without
    mt_state.dat) and motorCM (None without motor_positions.dat)."""
    cstep, n_a, n_b, n_s = read_timestep(os.path.join(directory, TIMESTEP))
    NT = int(engine.p["NT"])
    if n_a + n_b + n_s != NT:
        raise ValueError(f"timestep.dat counts {n_a}+{n_b}+{n_s} particles, the system has {NT}")
    fcs = engine.load_frame(os.path.join(directory, FRAME), react_box, tour)
    if fcs != cstep:
        raise ValueError(f"frame is at step {fcs}, timestep.dat at {cstep}")
    engine.setup(do_sort=(cstep % int(engine.p["sort_step"]) == 0), restart=True)
    mt = os.path.join(directory, MT_STATE)
    mp = os.path.join(directory, MOTOR_POSITIONS)
    return dict(cstep=cstep, streams=read_mt_state(mt) if os.path.exists(mt) else [],
                motor_cm=read_motor_positions(mp, int(engine.p["numdims"])) if os.path.exists(mp) else None)


def restart_from_h5(engine, directory: str, react_box, freqs, tour=None, h5name: str = "data.h5") -> dict:
    """Restore `engine` as restart_mod.f95:359-495 does from the reference's
    own files: the step and species counts of timestep.dat, the solvent of
    data.h5's frame cstep/Freq1 and the dimers of frame cstep/Freq4 (forces
    re-evaluated by the restart set-up), the streams of mt_state.dat and
    motorCM of motor_positions.dat.  freqs = (Freq1, Freq2, Freq3, Freq4)."""
    from . import h5data
    cstep, n_a, n_b, n_s = read_timestep(os.path.join(directory, TIMESTEP))
    p = engine.p
    if n_a + n_b + n_s != int(p["NT"]):
        raise ValueError(f"timestep.dat counts {n_a}+{n_b}+{n_s} particles, the system has {int(p['NT'])}")
    st = h5data.read_restart(os.path.join(directory, h5name), cstep, p, freqs)
    for k in ("fxd", "fyd", "fzd"):
        st[k] = np.zeros((int(p["numdims"]), 2), order="F")
    st["cstep"], st["react_box"] = cstep, react_box
    engine.set_state(st, tour=tour)
    engine.setup(do_sort=(cstep % int(p["sort_step"]) == 0), restart=True)
    mt = os.path.join(directory, MT_STATE)
    mp = os.path.join(directory, MOTOR_POSITIONS)
    return dict(cstep=cstep, streams=read_mt_state(mt) if os.path.exists(mt) else [],
                motor_cm=read_motor_positions(mp, int(p["numdims"])) if os.path.exists(mp) else None)
