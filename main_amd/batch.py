"""Batches of collective systems advanced together on one MI355X (see
mpcx_cob_* in include/mpcx.h): independent replicas with their own boxes,
parameters and Philox keys, one launch per kernel of an interval for all of
them, the normal intervals replayed from a HIP graph; FP64 or split-precision
replicas.  The tile size of the particle kernels (for split replicas it also
sets the cells per workgroup of the cell pass: each particle tile's owned
cells, about a tenth as many) and the slots of the split binning's hash
table are chosen by `autotune`, which times short runs.
"""
from __future__ import annotations

import ctypes as C

from ._lib import check, lib, track
from .collective import CollectiveEngine


class CollectiveBatch:
    def __init__(self, engines: list[CollectiveEngine], tile: int = 2048):
        self.engines = engines     # kept alive: the batch works on their device state
        arr = (C.c_void_p * len(engines))(*[e.h.value for e in engines])
        self.h = C.c_void_p()
        check(lib().mpcx_cob_create(C.byref(self.h), C.c_int32(len(engines)), arr, C.c_int32(tile)),
              "mpcx_cob_create")
        track(self, first=True)
        self.tile = tile
        self.bulk_slots = 0

    def close(self):
        if self.h:
            lib().mpcx_cob_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, ncoll: int, use_graph: bool = True):
        check(lib().mpcx_cob_run(self.h, C.c_int64(ncoll), C.c_int32(1 if use_graph else 0)), "mpcx_cob_run")

    def set_tile(self, tile: int):
        check(lib().mpcx_cob_set_tile(self.h, C.c_int32(tile)), "mpcx_cob_set_tile")
        self.tile = tile

    def info(self) -> dict:
        p, c, ms = C.c_int64(0), C.c_int64(0), C.c_double(0)
        check(lib().mpcx_cob_info(self.h, C.byref(p), C.byref(c), C.byref(ms)), "mpcx_cob_info")
        return dict(particles=p.value, cells=c.value, run_ms=ms.value)

    KERNELS = ("near_engine", "sb_bulk", "sb_cells_own", "sb_rotate", "sb_near")

    def kernel_ms(self) -> dict:
        """Mean kernel durations (ms) of the last split run without the graph."""
        import numpy as np
        from ._lib import ptr
        a = np.zeros(5)
        n = C.c_int64(0)
        check(lib().mpcx_cob_kernel_ms(self.h, ptr(a), C.byref(n)), "mpcx_cob_kernel_ms")
        d = dict(zip(self.KERNELS, a.tolist()))
        d["timed_intervals"] = n.value
        return d

    def body_average(self, n: int, nr=32, nz=64, rmax=12.0, zmin=-20.0, zmax=12.0):
        """Run n intervals; per replica the mean body-frame B counts (nz, nr)
        and the mean propulsion velocity over the n samples."""
        import numpy as np
        R = len(self.engines)
        counts = np.zeros((R, nz, nr), np.uint32)
        vsum = np.zeros(R)
        from ._lib import ptr
        check(lib().mpcx_cob_body_average(self.h, C.c_int64(n), C.c_int32(nr), C.c_int32(nz), C.c_double(rmax),
                                          C.c_double(zmin), C.c_double(zmax), ptr(counts), ptr(vsum)),
              "mpcx_cob_body_average")
        return counts.astype(np.float64) / max(n, 1), vsum / max(n, 1)

    def set_bulk_slots(self, slots: int):
        """Collision-cell slots of the split binning's hash table (1024, 2048,
        or 0 for the default by tile size)."""
        check(lib().mpcx_cob_set_bulk_slots(self.h, C.c_int32(slots)), "mpcx_cob_set_bulk_slots")
        self.bulk_slots = slots

    def set_inbox_cap(self, cap: int):
        """Entries per inbox of a split batch (0: the default by tile size)."""
        check(lib().mpcx_cob_set_inbox_cap(self.h, C.c_int32(cap)), "mpcx_cob_set_inbox_cap")

    def set_sort_step(self, steps: int):
        """Re-sort period of a split batch in MD steps (a multiple of MPCtstep)."""
        check(lib().mpcx_cob_set_sort_step(self.h, C.c_int32(steps)), "mpcx_cob_set_sort_step")
        self.sort_step = steps

    def tune_sort_period(self, periods=(10, 20, 40), use_graph: bool = True) -> dict:
        """Split batches: time one whole re-sort period (p intervals, one
        re-sort) at each period p in collisions and keep the fastest; returns
        {p: ms per interval}."""
        S = int(self.engines[0].p["MPCtstep"])
        times = {}
        for p in periods:
            self.set_sort_step(p * S)
            self.run(1, use_graph)   # capture / warm
            self.run(p, use_graph)
            times[p] = self.info()["run_ms"] / p
        best = min(times, key=times.get)
        self.set_sort_step(best * S)
        self.sort_period = best
        return times

    def autotune(self, tiles=(512, 1024, 2048, 4096, 8192), ncoll: int = 4, use_graph: bool = True,
                 slots=None) -> dict:
        """Time `ncoll` intervals at each tile size -- and, given slots, at
        each table size of the split binning -- advancing the systems, and
        keep the fastest; returns {tile: ms per interval}, or {(tile, slots):
        ms per interval}.  With re-sorts every k intervals, ncoll = k gives
        every setting one re-sort and the same spread of particle disorder."""
        times = {}
        for t in tiles:
            self.set_tile(t)
            for g in (slots or (None,)):
                if g is not None:
                    self.set_bulk_slots(g)
                self.run(1, use_graph)          # capture / warm
                self.run(ncoll, use_graph)
                times[t if g is None else (t, g)] = self.info()["run_ms"] / ncoll
        best = min(times, key=times.get)
        if isinstance(best, tuple):
            self.set_tile(best[0])
            self.set_bulk_slots(best[1])
        else:
            self.set_tile(best)
        return times
