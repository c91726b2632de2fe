"""Host side of the selkov_code path on the MI355X engine.

SelkovEngine mirrors the reference's time loop (Main_SelkovDimer.f90:125-149):
``run(n)`` advances n collision intervals, each MPCtstep calls of
Motion_SelkovDimer + ForcesMD_SelkovDimer, the last of which also does
Distribute, Affected_Cell and RMPC (and Sort every sort_step steps).  State
arrays have the reference's names, kinds (FLAG int64) and layouts.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ._lib import Mt95Stream, check, lib, ptr, require_gpu, track


class SkConfig(C.Structure):
    _fields_ = [("NT", C.c_int64), ("space", C.c_int64), ("MPCtstep", C.c_int32), ("sort_step", C.c_int32),
                ("rng_mode", C.c_int32), ("pad", C.c_int32), ("seed", C.c_uint64), ("box", C.c_double * 3)] + \
               [(k, C.c_double) for k in ("MDt", "MPCt", "Mf", "Mg", "Mi", "Mc", "Mn", "Dcn", "Rc", "Rn", "EfC",
                                          "EgC", "EiC", "EfN", "EgN", "EiN", "RPfg", "RPgf", "A0")] + \
               [("k", C.c_double * 6), ("sinrot", C.c_double), ("cosrot", C.c_double), ("sphere_guard", C.c_double)]


def make_config(p: dict, rng_mode: int = 0, seed: int = 8154543, sort_step: int = 10000,
                sphere_guard: float = 1.0) -> SkConfig:
    c = SkConfig()
    c.NT, c.space = int(p["NT"]), int(p.get("space", p["NT"]))
    c.MPCtstep, c.sort_step, c.rng_mode, c.seed = int(p["MPCtstep"]), int(sort_step), int(rng_mode), int(seed)
    for i in range(3):
        c.box[i] = float(p["box"][i])
    for k in ("MDt", "MPCt", "Mf", "Mg", "Mi", "Mc", "Mn", "Dcn", "Rc", "Rn", "EfC", "EgC", "EiC", "EfN", "EgN",
              "EiN", "RPfg", "RPgf", "A0"):
        setattr(c, k, float(p[k]))
    for i in range(6):
        c.k[i] = float(p["k%d" % (i + 1)])
    # Global_SelkovDimer.f90:327-328: the rotation angle is pi/2
    c.sinrot, c.cosrot = math.sin(math.pi / 2.0), math.cos(math.pi / 2.0)
    c.sphere_guard = sphere_guard
    return c


class SelkovEngine:
    """Device-resident selkov_code system (one GPU)."""

    def __init__(self, p: dict, rng_mode: int = 0, seed: int = 8154543, sort_step: int = 10000,
                 sphere_guard: float = 1.0):
        require_gpu()
        self.p = p
        self.cfg = make_config(p, rng_mode, seed, sort_step, sphere_guard)
        self.h = C.c_void_p()
        check(lib().mpcx_sk_create(C.byref(self.h), C.byref(self.cfg)), "mpcx_sk_create")
        track(self)
        self.rng_mode = rng_mode
        self.space = int(self.cfg.space)

    def close(self):
        if self.h:
            lib().mpcx_sk_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, s: dict, cstep: int = 0):
        f = lambda k: np.ascontiguousarray(s[k], dtype=np.float64)
        keep = [f(k) for k in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz")]
        flag = np.ascontiguousarray(s["flag"], dtype=np.int64)
        dm = np.asfortranarray(s["dm"], dtype=np.float64)
        check(lib().mpcx_sk_set_state(self.h, C.c_int64(cstep), C.c_int64(int(s["NT"])), *[ptr(a) for a in keep],
                                      ptr(flag), ptr(dm)), "mpcx_sk_set_state")

    def scalars(self) -> dict:
        sc = np.zeros(4, np.int64)
        tp = np.zeros(1)
        check(lib().mpcx_sk_get_state(self.h, *([None] * 10), None, ptr(sc), ptr(tp)), "mpcx_sk_get_state")
        return dict(cstep=int(sc[0]), NT=int(sc[1]), used=int(sc[2]), near_num=int(sc[3]), tpotential=float(tp[0]))

    def run(self, ncoll: int, stream: Mt95Stream | None = None) -> int:
        """Advance ncoll intervals; in rng_mode 0 the draws come from `stream`,
        which is advanced by exactly the number the run took."""
        if self.rng_mode != 0:
            check(lib().mpcx_sk_run(self.h, C.c_int64(ncoll), None, C.c_int64(0), None), "mpcx_sk_run")
            return 0
        if stream is None:
            raise ValueError("rng_mode 0 needs an mt95 stream")
        vol = int(np.prod([int(b) for b in self.p["box"]]))
        near = max(self.scalars()["near_num"], 1024)
        ub = ncoll * (3 + 4 * vol + int(self.p["MPCtstep"]) * 2 * near)
        saved_val, saved_cnt = stream.val.copy(), stream.cnt.value
        draws = stream.draws(ub)
        used = C.c_int64(0)
        check(lib().mpcx_sk_run(self.h, C.c_int64(ncoll), ptr(draws), C.c_int64(ub), C.byref(used)), "mpcx_sk_run")
        stream.val[:] = saved_val
        stream.cnt.value = saved_cnt
        stream.draws(int(used.value))
        return int(used.value)

    def get_state(self) -> dict:
        n = self.space
        out = {k: np.empty(n) for k in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz")}
        out["flag"] = np.empty(n, np.int64)
        dm = np.zeros((2, 12), order="F")
        sc = np.zeros(4, np.int64)
        tp = np.zeros(1)
        check(lib().mpcx_sk_get_state(self.h, *[ptr(out[k]) for k in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy",
                                                                      "fz")],
                                      ptr(out["flag"]), ptr(dm), ptr(sc), ptr(tp)), "mpcx_sk_get_state")
        out["dm"] = dm
        out["cstep"], out["NT"], out["used"], out["near_num"] = (int(v) for v in sc)
        out["tpotential"] = float(tp[0])
        return out

    def get_cells(self) -> dict:
        vol = int(np.prod([int(b) for b in self.p["box"]]))
        rec = np.empty((vol, 8))
        cell = np.empty(self.space, np.int32)
        census = np.empty((3, vol), np.int32)
        unaff = np.empty(vol, np.uint8)
        check(lib().mpcx_sk_get_cells(self.h, ptr(rec), ptr(cell), ptr(census), ptr(unaff)), "mpcx_sk_get_cells")
        return dict(vxcm=rec[:, 0].copy(), vycm=rec[:, 1].copy(), vzcm=rec[:, 2].copy(), uvx=rec[:, 3].copy(),
                    uvy=rec[:, 4].copy(), uvz=rec[:, 5].copy(), tcs=rec[:, 6].astype(np.int64),
                    ncn=cell.astype(np.int64) + 1, tF=census[0].astype(np.int64), tG=census[1].astype(np.int64),
                    tI=census[2].astype(np.int64), affected=(unaff == 0))

    OBSERVABLES = ("n", "nF", "nG", "nI", "psol_x", "psol_y", "psol_z", "p_x", "p_y", "p_z", "ke_sol", "ke_dim",
                   "T_sol", "T_dim", "E_tot", "v_prop")

    def observables(self) -> dict:
        a = np.zeros(16)
        check(lib().mpcx_sk_observables(self.h, ptr(a)), "mpcx_sk_observables")
        return dict(zip(self.OBSERVABLES, a.tolist()))

    def shells(self) -> np.ndarray:
        """F, G, I counts (3, 8) in the shells of Output_Concentration_Catalytic.f90."""
        a = np.zeros((3, 8), np.int64)
        check(lib().mpcx_sk_shells(self.h, ptr(a)), "mpcx_sk_shells")
        return a

    def body_density(self, nr=32, nz=64, rmax=16.0, zmin=-24.0, zmax=16.0) -> np.ndarray:
        a = np.zeros((3, nz, nr), np.uint32)
        check(lib().mpcx_sk_body_density(self.h, C.c_int32(nr), C.c_int32(nz), C.c_double(rmax), C.c_double(zmin),
                                         C.c_double(zmax), ptr(a)), "mpcx_sk_body_density")
        return a

    def set_overlap(self, on: bool):
        """With device draws (rng_mode 1), stream and bin the far particles
        on a second stream beside the near engine (the default); the states
        are bitwise those of the one-stream order."""
        check(lib().mpcx_sk_set_overlap(self.h, C.c_int32(1 if on else 0)), "mpcx_sk_set_overlap")

    KERNELS = ("sk_near_engine", "sk_bulk", "sk_collide", "sk_rotate")

    def kernel_ms(self) -> dict:
        a = np.zeros(4)
        r = C.c_double(0)
        check(lib().mpcx_sk_kernel_ms(self.h, ptr(a), C.byref(r)), "mpcx_sk_kernel_ms")
        d = dict(zip(self.KERNELS, a.tolist()))
        d["run_ms"] = r.value
        return d
