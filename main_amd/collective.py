"""Host side of the collective_code path on the MI355X engine.

CollectiveEngine mirrors the reference's time loop (main_pattern.f95:170-240):
``setup()`` is the set-up of :170-188 and ``run(n)`` performs n collision
intervals of :202-225 (MPCtstep calls of evolve, then distribute,
mpc_rotation and reaction_simple).  The state arrays it takes and returns
have the reference's names, kinds and layouts.  The per-call drop-ins of the
individual module procedures live in ``main_amd.dropin``.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

from ._lib import MtmodStream, check, lib, ptr, require_gpu, track

MAXM = 8            # force slots stored per inside entry by the engine (one per list of up to four dimers)
MAX_INSIDE = 16384  # inside-list capacity of the near engine


class CoConfig(C.Structure):
    _fields_ = [("NT", C.c_int64), ("numdims", C.c_int64), ("vol", C.c_int64), ("mpcnumcell", C.c_int64),
                ("rxnnumcell", C.c_int64), ("max_neighbour", C.c_int64),
                ("MPCtstep", C.c_int32), ("sort_step", C.c_int32), ("react_dir", C.c_int32),
                ("rng_mode", C.c_int32), ("seed", C.c_uint64), ("box", C.c_double * 3)] + \
               [(k, C.c_double) for k in ("MDt", "Ma", "Mb", "Ms", "Mc", "Mn", "Dcn", "cutoff", "Rc", "Rn",
                                          "EaC", "EbC", "EsC", "EaN", "EbN", "EsN", "Rw", "Ew", "Rm_delta",
                                          "Em", "RPab", "RPba", "skinrad", "rotang", "a0", "MPCt", "k1",
                                          "sinrot", "cosrot", "wall_w6")] + \
               [("reaction_p", C.c_double * 101), ("sphere_guard", C.c_double)]


def reaction_table(k1: float, MPCt: float) -> np.ndarray:
    """prob1(n) = 1 - dexp(-a1*MPCt), a1 = k1*n (reaction_mod.f95:45-54), host libm."""
    t = np.zeros(101)
    for n in range(101):
        a1 = 0.0 if n == 0 else k1 * float(n)
        t[n] = 1.0 - math.exp(-(a1 * MPCt))
    return t


def make_config(p: dict, rng_mode: int = 0, seed: int = 4357787, sphere_guard: float = 0.3) -> CoConfig:
    c = CoConfig()
    for k in ("NT", "numdims", "vol", "mpcnumcell", "rxnnumcell"):
        setattr(c, k, int(p[k]))
    c.max_neighbour = int(p.get("max_nb", 30000))
    c.MPCtstep, c.sort_step, c.react_dir = int(p["MPCtstep"]), int(p["sort_step"]), int(p["react_dir"])
    c.rng_mode, c.seed = int(rng_mode), int(seed)
    for i in range(3):
        c.box[i] = float(p["box"][i])
    for k in ("MDt", "Ma", "Mb", "Ms", "Mc", "Mn", "Dcn", "cutoff", "Rc", "Rn", "EaC", "EbC", "EsC", "EaN",
              "EbN", "EsN", "Rw", "Ew", "Rm_delta", "Em", "RPab", "RPba", "skinrad", "rotang", "a0", "MPCt", "k1"):
        setattr(c, k, float(p[k]))
    c.sinrot = math.sin(p["rotang"])
    c.cosrot = math.cos(p["rotang"])
    c.wall_w6 = 3.0 ** float(np.float32(1.0) / np.float32(6.0))
    tab = reaction_table(p["k1"], p["MPCt"])
    for i in range(101):
        c.reaction_p[i] = tab[i]
    c.sphere_guard = sphere_guard
    return c


def draws_per_collision(p: dict) -> int:
    """grnd() calls of one collision at one thread: 3 shift + 2 per MPC cell +
    1 per reaction cell (mpc_rotation_mod.f95:102-107,187-189; reaction_mod.f95:57)."""
    return 3 + 2 * int(p["mpcnumcell"]) + int(p["rxnnumcell"])


class CollectiveEngine:
    """Device-resident collective_code system (one GPU)."""

    def __init__(self, p: dict, rng_mode: int = 0, seed: int = 4357787, sphere_guard: float = 0.3,
                 near_mode: int | None = None, near_workgroups: int | None = None, split: bool = False,
                 overlap: bool = True):
        """near_mode (two or more dimers): 0 one near-engine workgroup for the
        system (up to four dimers), 1 one per cluster of dimers, 2 the near
        set over cooperative workgroups (default above two dimers); see
        mpcx_co_set_near_mode.  overlap: far particles in one flight beside
        the near engine (False: the bitwise event replay after it); see
        mpcx_co_set_overlap."""
        require_gpu()
        self.p = p
        self.cfg = make_config(p, rng_mode, seed, sphere_guard)
        self.h = C.c_void_p()
        check(lib().mpcx_co_create(C.byref(self.h), C.byref(self.cfg)), "mpcx_co_create")
        track(self)
        self.rng_mode = rng_mode
        if near_mode is not None:
            check(lib().mpcx_co_set_near_mode(self.h, C.c_int32(int(near_mode))), "mpcx_co_set_near_mode")
        if split:
            check(lib().mpcx_co_set_precision(self.h, C.c_int32(1)), "mpcx_co_set_precision")
        self.split = split
        # co_far beside the near engine: one-workgroup mode (the default up to two dimers), FP64 storage
        self._ov_mode = near_mode if near_mode is not None else (0 if int(p["numdims"]) <= 2 else 2)
        self.set_overlap(overlap)
        if near_workgroups is not None:
            check(lib().mpcx_co_set_near_workgroups(self.h, C.c_int32(int(near_workgroups))),
                  "mpcx_co_set_near_workgroups")

    def set_overlap(self, on: bool):
        """Far particles in one flight beside the near engine (True) or the
        bitwise event replay after it (False), from the next run on."""
        check(lib().mpcx_co_set_overlap(self.h, C.c_int32(1 if on else 0)), "mpcx_co_set_overlap")
        self.overlap = bool(on) and self._ov_mode == 0

    def close(self):
        if self.h:
            lib().mpcx_co_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, s: dict, tour=None):
        f = lambda k: np.ascontiguousarray(s[k], dtype=np.float64)
        dims = {k: np.asfortranarray(s[k], dtype=np.float64) for k in
                ("xd", "yd", "zd", "vxd", "vyd", "vzd", "fxd", "fyd", "fzd")}
        self._keep = [f(k) for k in ("x", "y", "z", "vx", "vy", "vz")]
        flag = np.ascontiguousarray(s["flag"], dtype=np.int32)
        rb = np.ascontiguousarray(s["react_box"], dtype=np.int32)
        t = None if tour is None else np.ascontiguousarray(tour, dtype=np.int64)
        check(lib().mpcx_co_set_state(self.h, C.c_int32(int(s.get("cstep", 0))), *[ptr(a) for a in self._keep],
                                      ptr(flag), *[ptr(dims[k]) for k in ("xd", "yd", "zd", "vxd", "vyd", "vzd",
                                                                           "fxd", "fyd", "fzd")],
                                      ptr(rb), ptr(t)), "mpcx_co_set_state")

    def set_dimers(self, s: dict, tour=None, cstep: int = 0):
        """Dimers, reactive grid and tour only; the solvent then comes from init_solvent."""
        dims = {k: np.asfortranarray(s[k], dtype=np.float64) for k in
                ("xd", "yd", "zd", "vxd", "vyd", "vzd", "fxd", "fyd", "fzd")}
        self._keep = list(dims.values())
        rb = np.ascontiguousarray(s["react_box"], dtype=np.int32)
        t = None if tour is None else np.ascontiguousarray(tour, dtype=np.int64)
        self._keep += [rb, t]
        null = None
        check(lib().mpcx_co_set_state(self.h, C.c_int32(cstep), *([null] * 7),
                                      *[ptr(dims[k]) for k in ("xd", "yd", "zd", "vxd", "vyd", "vzd",
                                                               "fxd", "fyd", "fzd")],
                                      ptr(rb), ptr(t)), "mpcx_co_set_state")

    def init_solvent(self, seed: int):
        """Solvent of initial_mod.f95:122-345 generated on the device (Philox draws)."""
        p = self.p
        check(lib().mpcx_co_init_solvent(self.h, C.c_int64(int(p["Na0"])), C.c_int64(int(p["Nb0"])),
                                         C.c_double(float(p["T"])), C.c_uint64(int(seed))), "mpcx_co_init_solvent")

    def setup(self, do_sort: bool = True, restart: bool = False):
        """Set-up of main_pattern.f95:170-188 (zero forces, as after
        initial_mod.f95:348-353) or, with restart, of restart_mod.f95:416-495
        (forces evaluated from the uploaded positions)."""
        if restart:
            check(lib().mpcx_co_setup_restart(self.h, C.c_int32(1 if do_sort else 0)), "mpcx_co_setup_restart")
        else:
            check(lib().mpcx_co_setup(self.h, C.c_int32(1 if do_sort else 0)), "mpcx_co_setup")

    def save_frame(self, path: str):
        """Frame of the system (solvent and dimer positions, velocities, species)
        at a collision step; see mpcx_co_save_frame."""
        check(lib().mpcx_co_save_frame(self.h, os.fsencode(path)), "mpcx_co_save_frame")

    def load_frame(self, path: str, react_box, tour=None) -> int:
        """Upload a frame written by save_frame; returns its collision step."""
        rb = np.ascontiguousarray(react_box, dtype=np.int32)
        t = None if tour is None else np.ascontiguousarray(tour, dtype=np.int64)
        cs = C.c_int32(0)
        check(lib().mpcx_co_load_frame(self.h, os.fsencode(path), ptr(rb), ptr(t), C.byref(cs)),
              "mpcx_co_load_frame")
        return int(cs.value)

    def run(self, ncoll: int, draws: np.ndarray | None = None):
        d = None if draws is None else np.ascontiguousarray(draws, dtype=np.float64)
        if self.rng_mode == 0:
            need = ncoll * draws_per_collision(self.p)
            if d is None or d.size < need:
                raise ValueError(f"rng_mode 0 needs {need} host draws")
        check(lib().mpcx_co_run(self.h, C.c_int64(ncoll), ptr(d)), "mpcx_co_run")

    def get_state(self) -> dict:
        NT, nd = self.p["NT"], self.p["numdims"]
        out = {k: np.empty(NT) for k in ("x", "y", "z", "vx", "vy", "vz")}
        out["flag"] = np.empty(NT, np.int32)
        dim = np.empty(12 * 2 * nd)
        sc = np.zeros(4, np.int32)
        check(lib().mpcx_co_get_state(self.h, *[ptr(out[k]) for k in ("x", "y", "z", "vx", "vy", "vz")],
                                      ptr(out["flag"]), ptr(dim), ptr(sc)), "mpcx_co_get_state")
        dim = dim.reshape(12, 2 * nd)
        for f, k in enumerate(("xd", "yd", "zd", "vxd", "vyd", "vzd", "fxd", "fyd", "fzd", "qx", "qy", "qz")):
            out[k] = np.asfortranarray(dim[f].reshape(2, nd).T)   # s = k + i*nd -> (k, i)
        out["cstep"], out["time_count"], out["time_penetrate"], out["in_num"] = (int(v) for v in sc)
        return out

    def scalars(self) -> dict:
        """cstep, time_count, time_penetrate and the inside-list length."""
        sc = np.zeros(4, np.int32)
        check(lib().mpcx_co_get_state(self.h, *([None] * 8), ptr(sc)), "mpcx_co_get_state")
        return dict(cstep=int(sc[0]), time_count=int(sc[1]), time_penetrate=int(sc[2]), in_num=int(sc[3]))

    def get_lists(self) -> dict:
        """Inside list, neighbour lists (reference layout, 1-based) and Fpair."""
        nd, NT, mnb = self.p["numdims"], self.p["NT"], int(self.p.get("max_nb", 30000))
        cap = MAX_INSIDE * max(1, nd)
        lst = np.zeros(cap, np.int32)
        nb = np.zeros((nd, 2, mnb), np.int64, order="F")
        nbnum = np.zeros((nd, 2), np.int64, order="F")
        fpx = np.zeros((nd, 2, mnb), order="F")
        fpy, fpz, potn = fpx.copy(order="F"), fpx.copy(order="F"), fpx.copy(order="F")
        n = C.c_int64(0)
        check(lib().mpcx_co_get_lists(self.h, C.c_int64(mnb), ptr(nb), ptr(nbnum), ptr(fpx), ptr(fpy), ptr(fpz),
                                      ptr(potn), ptr(lst), C.c_int64(cap), C.byref(n)), "mpcx_co_get_lists")
        inside = lst[:n.value].astype(np.int64) + 1
        outside = np.setdiff1d(np.arange(1, NT + 1, dtype=np.int64), inside, assume_unique=True)
        return dict(in_list=inside, out_list=outside, nb=nb, nbnum=nbnum, fpx=fpx, fpy=fpy, fpz=fpz, pot=potn)

    def get_cells(self) -> dict:
        """The last collision's cell records (by collision cell), each
        particle's collision cell (1-based) and, in FP64 storage, the census
        (split storage clears its census after each collision: tA, tB, tS are
        None there)."""
        mc, NT, vol = self.p["mpcnumcell"], self.p["NT"], self.p["vol"]
        rec = np.empty((mc, 8))
        cell = np.empty(NT, np.int32)
        census = None if self.split else np.empty((3, vol), np.int32)
        tp = np.zeros(1)
        check(lib().mpcx_co_get_cells(self.h, ptr(rec), ptr(cell), None if census is None else ptr(census), ptr(tp)),
              "mpcx_co_get_cells")
        cen = (None, None, None) if census is None else tuple(census[i].astype(np.int64) for i in range(3))
        return dict(vxcm=rec[:, 0].copy(), vycm=rec[:, 1].copy(), vzcm=rec[:, 2].copy(), uvx=rec[:, 3].copy(),
                    uvy=rec[:, 4].copy(), uvz=rec[:, 5].copy(), tcs=rec[:, 6].copy(),
                    ncn=cell.astype(np.int64) + 1, tA=cen[0], tB=cen[1], tS=cen[2], tpotential=float(tp[0]))

    def tpotential(self) -> float:
        """TPotential of the last force evaluation (set-up or MD step)."""
        tp = np.zeros(1)
        check(lib().mpcx_co_get_cells(self.h, None, None, None, ptr(tp)), "mpcx_co_get_cells")
        return float(tp[0])

    def near_profile(self) -> dict:
        a = np.zeros(10, np.uint64)
        check(lib().mpcx_co_near_profile(self.h, ptr(a)), "mpcx_co_near_profile")
        us = lambda v: float(v) / 100.0   # 100 MHz ticks -> microseconds
        n = max(int(a[7]), 1)
        return dict(launches=int(a[7]), total_us=us(a[0]) / n, rebuild_us=us(a[1]) / n,
                    particle_us=us(a[2]) / n, serial_us=us(a[3]) / n, rebuilds=int(a[4]) / n,
                    near_num=int(a[5]), inside_num=int(a[6]), gather_us=us(a[8]) / n, end_us=us(a[9]) / n)

    OBSERVABLES = ("n", "nA", "nB", "nS", "psol_x", "psol_y", "psol_z", "p_x", "p_y", "p_z", "ke_sol", "ke_dim",
                   "T_sol", "T_dim", "E_tot", "v_prop")

    def observables(self) -> dict:
        """System observables of output_mod.f95:681-805 and the propulsion velocity."""
        a = np.zeros(16)
        check(lib().mpcx_co_observables(self.h, ptr(a)), "mpcx_co_observables")
        return dict(zip(self.OBSERVABLES, a.tolist()))

    def conc_slabs(self, x_slab, y_slab, z_slab, half_width, width=0.5) -> dict:
        """A/B counts of the concentration slabs (output_mod.f95:486-568)."""
        L = [int(b) for b in self.p["box"]]
        nx, ny, nz = (int(v / width) for v in L)
        sizes = [("A1d", (nx,)), ("B1d", (nx,)), ("Axy", (nx, ny)), ("Bxy", (nx, ny)), ("Axz", (nx, nz)),
                 ("Bxz", (nx, nz)), ("Ayz", (ny, nz)), ("Byz", (ny, nz))]
        buf = np.zeros(sum(int(np.prod(sh)) for _, sh in sizes), np.int32)
        check(lib().mpcx_co_conc_slabs(self.h, C.c_double(x_slab), C.c_double(y_slab), C.c_double(z_slab),
                                       C.c_double(half_width), C.c_double(width), C.c_double(width),
                                       C.c_double(width), ptr(buf)), "mpcx_co_conc_slabs")
        out, o = {}, 0
        for k, sh in sizes:
            n = int(np.prod(sh))
            out[k] = buf[o:o + n].reshape(sh, order="F")
            o += n
        return out

    def body_density(self, nr=32, nz=64, rmax=16.0, zmin=-24.0, zmax=16.0, dimer=0) -> np.ndarray:
        """Counts (3, nz, nr) per species in the body frame of a dimer."""
        a = np.zeros((3, nz, nr), np.uint32)
        check(lib().mpcx_co_body_density(self.h, C.c_int32(dimer), C.c_int32(nr), C.c_int32(nz), C.c_double(rmax),
                                         C.c_double(zmin), C.c_double(zmax), ptr(a)), "mpcx_co_body_density")
        return a

    KERNELS = ("co_near_engine", "co_bulk", "co_collide", "co_rotate", "co_scatter")

    def kernel_ms(self) -> dict:
        """Mean duration (ms) of each interval kernel over the last run (HIP events)."""
        a = np.zeros(5)
        n = C.c_int64(0)
        check(lib().mpcx_co_kernel_ms(self.h, ptr(a), C.byref(n)), "mpcx_co_kernel_ms")
        if self.split:   # the split interval's kernels (mpcx_cob_kernel_ms)
            names = ["near_engine", "sb_bulk", "sb_cells_own", "sb_rotate", "sb_near"]
        else:
            names = [("co_far" if self.overlap and k == "co_bulk" else k) for k in self.KERNELS]
        return dict(zip(names, a.tolist()), intervals=int(n.value))

    def timing(self):
        r, b, n = C.c_double(), C.c_double(), C.c_double()
        check(lib().mpcx_co_timing(self.h, C.byref(r), C.byref(b), C.byref(n)), "mpcx_co_timing")
        return r.value, b.value, n.value


def host_draws(stream: MtmodStream, p: dict, ncoll: int) -> np.ndarray:
    """The grnd() draws of ncoll collisions, in reference order."""
    return stream.draws(ncoll * draws_per_collision(p))
