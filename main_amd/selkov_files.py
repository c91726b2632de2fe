"""Text records and restart files of selkov_code for the MI355X engine.

The reference writes its output with fixed FORMAT statements
(Output_SelkovDimer.f90) and restarts from the same files (Restart.f90,
Main_SelkovDimer.f90:108-158).  The records are formatted by the engine
library (mpcx_format_record), byte for byte as the FORMAT statements:

* dimer records ``Dimer{Pos,Vel,For}{C,N}.dat``: FORMAT 15 (I8,1X,3(E23.16,1X)), :26;
* solvent dumps ``Solvent{F,G,I}{Pos,Vel,For}<cstep>.dat``: FORMAT 16 (3(ES23.16,1X),I8), :139;
* ``DotVZ.dat``: FORMAT 18 (I8,4(1X,ES23.16)), :184-186;
* ``Concentration.dat``: FORMAT 19 (I8,10(1X,ES23.16)), :237-239;
* ``EnergyMomentum.dat``: FORMAT 17 (I8,8(1X,ES23.16)), :303-305;
* ``timestep.dat``: FORMAT 23 (I9), Main_SelkovDimer.f90:152-158.

Every file is opened with STATUS='replace' in the reference, so each holds
the record of the last output step.  E23.16 keeps 16 significant digits and
ES23.16 17, so a restart gets the solvent back exactly and the dimer rounded
to 16 digits, as the reference's own restart does.  The sums of the energy/momentum record
are taken on the host in particle order (cumulative sums), as the reference
adds them; the temperatures use its single-precision constant 2./3.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

from ._lib import check, lib, ptr

FMT15 = "I8 X1 E23.16 X1 E23.16 X1 E23.16 X1"
FMT16 = "S23.16 X1 S23.16 X1 S23.16 X1 I8"
FMT17 = "I8" + " X1 S23.16" * 8
FMT18 = "I8" + " X1 S23.16" * 4
FMT19 = "I8" + " X1 S23.16" * 10
FMT23 = "I9"
TWO_THIRDS = float(np.float32(2.0) / np.float32(3.0))   # the default-real literal 2./3.
SPECIES = (("F", 0), ("G", 1), ("I", 3))


def format_record(spec: str, values) -> str:
    """One record of the fixed-form edit descriptors in spec (see mpcx_format_record)."""
    v = np.ascontiguousarray(np.asarray(values, dtype=np.float64).ravel())
    buf = C.create_string_buffer(64 * (v.size + 2) + 64)
    check(lib().mpcx_format_record(spec.encode(), ptr(v), C.c_int32(v.size), buf, C.c_int64(len(buf))),
          "mpcx_format_record")
    return buf.value.decode()


def _lines(spec: str, rows) -> str:
    return "".join(format_record(spec, r) + "\n" for r in rows)


def dump_name(stem: str, cstep: int) -> str:
    """File name of Output_SelkovDimer.f90:89-129: stem, cstep in I2..I8 by its
    decade ((0,100) -> I2, [100,1000) -> I3, ...), '.dat'."""
    if not 0 < cstep < 100_000_000:
        raise ValueError("the reference names dumps for 0 < cstep < 10**8 only")
    w = max(2, len(str(cstep)))
    return f"{stem}{cstep:>{w}d}.dat"


def restart_dump_name(stem: str, pstep: int) -> str:
    """File name Restart.f90:10-80 opens: the width chosen by pstep+1, the value pstep."""
    cs = pstep + 1
    if not 0 < cs < 100_000_000:
        raise ValueError("step outside the reference's naming ranges")
    w = max(2, len(str(cs)))
    return f"{stem}{pstep:>{w}d}.dat"


# ------------------------------------------------------------------ records
def dimer_records(state: dict, cstep: int) -> dict[str, str]:
    """Output_SelkovDimer.f90:7-28: the six dimer files (one record each)."""
    dm = np.asarray(state["dm"])
    out = {}
    for name, cols in (("Pos", (0, 1, 2)), ("Vel", (3, 4, 5)), ("For", (6, 7, 8))):
        for row, sph in ((0, "C"), (1, "N")):
            out[f"Dimer{name}{sph}.dat"] = format_record(FMT15, [cstep, *dm[row, list(cols)]]) + "\n"
    return out


def solvent_dumps(state: dict, cstep: int) -> dict[str, str]:
    """Output_SelkovDimer.f90:130-143: per species F (flag 0), G (1) and I (3),
    positions, velocities and forces with the 1-based particle index."""
    NT = int(state["NT"])
    flag = np.asarray(state["flag"])[:NT]
    idx = np.arange(1, NT + 1)
    out = {}
    for sp, f in SPECIES:
        sel = flag == f
        for name, keys in (("Pos", ("x", "y", "z")), ("Vel", ("vx", "vy", "vz")), ("For", ("fx", "fy", "fz"))):
            cols = np.stack([np.asarray(state[k])[:NT][sel] for k in keys] + [idx[sel].astype(np.float64)], axis=1)
            out[dump_name(f"Solvent{sp}{name}", cstep)] = _lines(FMT16, cols)
    return out


def dot_vz(state: dict, p: dict) -> tuple[float, float, float, float]:
    """DotVZ and the unit vector of Output_SelkovDimer.f90:147-181, the
    reference's operations in its order."""
    dm = np.asarray(state["dm"])
    box = [float(b) for b in p["box"]]
    d = []
    for a in range(3):
        v = dm[0, a] - dm[1, a]
        h = 0.5 * box[a]
        if v > h:
            v = v - box[a]
        elif v < -h:
            v = v + box[a]
        d.append(v)
    dxyz = np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
    nuc = [d[a] / dxyz for a in range(3)]
    Mc, Mn = float(p["Mc"]), float(p["Mn"])
    vav = [(Mc * dm[0, 3 + a] + Mn * dm[1, 3 + a]) / (Mn + Mc) for a in range(3)]
    dot = vav[0] * nuc[0] + vav[1] * nuc[1] + vav[2] * nuc[2]
    return float(dot), float(nuc[0]), float(nuc[1]), float(nuc[2])


def _seq_sum(a: np.ndarray) -> float:
    """Left-to-right sum from 0, as a Fortran DO loop adds."""
    return float(np.cumsum(np.concatenate([[0.0], a]))[-1]) if a.size else 0.0


def energy_momentum(state: dict, p: dict) -> list[float]:
    """sumPX, sumPY, sumPZ, Ekin, Epot, Etot, tempsol, tempdim of
    Output_SelkovDimer.f90:243-305 (the dimer velocity and the species totals
    taken at this step)."""
    NT = int(state["NT"])
    flag = np.asarray(state["flag"])[:NT]
    m = np.zeros(NT)
    for key, f in (("Mf", 0), ("Mg", 1), ("Mi", 3)):
        m[flag == f] = float(p[key])
    sel = (flag == 0) | (flag == 1) | (flag == 3)
    v = [np.asarray(state[k])[:NT][sel] for k in ("vx", "vy", "vz")]
    P = [m[sel] * v[a] for a in range(3)]
    psol = [_seq_sum(P[a]) for a in range(3)]
    Mc, Mn = float(p["Mc"]), float(p["Mn"])
    dm = np.asarray(state["dm"])
    vav = [(Mc * dm[0, 3 + a] + Mn * dm[1, 3 + a]) / (Mn + Mc) for a in range(3)]
    Pd = [(Mc + Mn) * vav[a] for a in range(3)]
    s = [psol[a] + Pd[a] for a in range(3)]
    ke_sol = _seq_sum(0.5 * ((v[0] * P[0] + v[1] * P[1]) + v[2] * P[2]))
    ke_dim = 0.0 + (0.5 * (1.0 / (Mc + Mn))) * ((Pd[0] * Pd[0] + Pd[1] * Pd[1]) + Pd[2] * Pd[2])
    ekin = ke_dim + ke_sol
    epot = float(state["tpotential"])
    n = float(np.count_nonzero(flag == 0)) + float(np.count_nonzero(flag == 1)) + float(np.count_nonzero(flag == 3))
    return [*s, ekin, epot, ekin + epot, (ke_sol / n) * TWO_THIRDS, ke_dim * TWO_THIRDS]


def concentration(state: dict, cells: dict, p: dict) -> list[float]:
    """TotF..TBulkCells of Output_SelkovDimer.f90:190-239 (bulk = cells the
    catalytic sphere does not affect)."""
    NT = int(state["NT"])
    flag = np.asarray(state["flag"])[:NT]
    vol = int(np.prod([int(b) for b in p["box"]]))
    tot = [float(np.count_nonzero(flag == f)) for _, f in SPECIES]
    bulk = ~np.asarray(cells["affected"])[:vol]
    sums = [_seq_sum(np.asarray(cells[k])[:vol][bulk].astype(np.float64)) for k in ("tF", "tG", "tI")]
    nb = float(np.count_nonzero(bulk))
    return [*tot, *[t / vol for t in tot], *[x / nb for x in sums], nb]


# ------------------------------------------------------------------- output
def write_files(directory: str, files: dict[str, str]):
    os.makedirs(directory, exist_ok=True)
    for name, text in files.items():
        with open(os.path.join(directory, name), "w") as f:
            f.write(text)


def write_timestep(directory: str, cstep: int, counts: tuple[int, int, int]):
    """timestep.dat of Main_SelkovDimer.f90:152-158 (FORMAT 23, I9)."""
    write_files(directory, {"timestep.dat": _lines(FMT23, [[cstep], [counts[0]], [counts[1]], [counts[2]]])})


def output_records(engine, p: dict) -> dict[str, str]:
    """The record files of Output_SelkovDimer.f90 for the engine's state now
    (the reference writes each at its own cadence Freq1..Freq5): dimer
    records, solvent dumps, DotVZ, Concentration and EnergyMomentum."""
    s = engine.get_state()
    cstep = int(s["cstep"])
    files = dimer_records(s, cstep)
    files.update(solvent_dumps(s, cstep))
    files["DotVZ.dat"] = format_record(FMT18, [cstep, *dot_vz(s, p)]) + "\n"
    files["Concentration.dat"] = format_record(FMT19, [cstep, *concentration(s, engine.get_cells(), p)]) + "\n"
    files["EnergyMomentum.dat"] = format_record(FMT17, [cstep, *energy_momentum(s, p)]) + "\n"
    return files


def save_restart(engine, directory: str) -> int:
    """Write what Restart.f90 reads (dimer records, solvent dumps, timestep.dat)
    for the engine's state at a collision step; returns cstep."""
    s = engine.get_state()
    cstep = int(s["cstep"])
    files = dimer_records(s, cstep)
    files.update(solvent_dumps(s, cstep))
    write_files(directory, files)
    flag = np.asarray(s["flag"])[: int(s["NT"])]
    write_timestep(directory, cstep, tuple(int(np.count_nonzero(flag == f)) for _, f in SPECIES))
    return cstep


# -------------------------------------------------------------------- input
_EXP3 = re.compile(r"(?<=[0-9.])([+-]\d{3})(?!\d)")


def _numbers(text: str) -> np.ndarray:
    """Whitespace-separated values of fixed-form records (an exponent above
    99 is written without its letter: 0.1000000000000000+100)."""
    return np.array(_EXP3.sub(r"E\1", text).split(), dtype=np.float64)


def read_timestep(directory: str) -> tuple[int, int, int, int]:
    """cstep, iTotF, iTotG, iTotI as Main_SelkovDimer.f90:108-113 reads them."""
    with open(os.path.join(directory, "timestep.dat")) as f:
        v = [int(f.readline()[:9]) for _ in range(4)]
    return v[0], v[1], v[2], v[3]


def read_restart(directory: str, p: dict) -> dict:
    """The state Restart.f90 builds: particles ordered F, G, I with flags 0, 1,
    3, positions, velocities and forces from the dumps, the dimer from the
    dimer records (the record whose step is cstep, else the last)."""
    cstep, nf, ng, ni = read_timestep(directory)
    NT = nf + ng + ni
    st = {k: np.zeros(NT) for k in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz")}
    st["flag"] = np.zeros(NT, np.int64)
    o = 0
    for (sp, f), n in zip(SPECIES, (nf, ng, ni)):
        for name, keys in (("Pos", ("x", "y", "z")), ("Vel", ("vx", "vy", "vz")), ("For", ("fx", "fy", "fz"))):
            path = os.path.join(directory, restart_dump_name(f"Solvent{sp}{name}", cstep))
            if not os.path.exists(path):
                path = os.path.join(directory, dump_name(f"Solvent{sp}{name}", cstep))
            with open(path) as fh:
                a = _numbers(fh.read()).reshape(-1, 4)
            if a.shape[0] < n:
                raise ValueError(f"{path}: {a.shape[0]} records, timestep.dat counts {n}")
            for c, k in enumerate(keys):
                st[k][o:o + n] = a[:n, c]
        st["flag"][o:o + n] = f
        o += n
    dm = np.zeros((2, 12), order="F")
    for name, cols in (("Pos", (0, 1, 2)), ("Vel", (3, 4, 5)), ("For", (6, 7, 8))):
        for row, sph in ((0, "C"), (1, "N")):
            with open(os.path.join(directory, f"Dimer{name}{sph}.dat")) as fh:
                recs = _numbers(fh.read()).reshape(-1, 4)
            hit = np.nonzero(recs[:, 0] == cstep)[0]
            r = recs[hit[-1]] if hit.size else recs[-1]
            dm[row, list(cols)] = r[1:]
    st["dm"] = dm
    st["NT"] = NT
    st["cstep"] = cstep
    return st
