"""The no-allocation entry point mpcx_inout_particle_list_noalloc against the
allocating mpcx_inout_particle_list on the same particles: the same inside and
outside lists, bit for bit, on two calls in a row (the second with fewer
particles, so it reuses the fixed scratch array a larger call filled), each
appending after entries already in the lists; sizes beyond its limits are
rejected before any GPU work."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from main_amd import dropin as D
from main_amd._lib import MpcxError, lib
from tests.parity_cases import build_collective, small_collective_params

pytestmark = pytest.mark.gpu


def _system(numdims):
    p = small_collective_params(numdims=numdims, Lx=20.0, Ly=20.0, Lz=24.0)
    return p, build_collective(p, seed=17, flags_mix=True)


def _allocating(p, s, NT, pre):
    on, inn = np.array([pre], np.int64), np.array([pre], np.int64)
    ol, il = np.full(NT + pre, -7, np.int64), np.full(NT + pre, -7, np.int64)
    D.inout_particle_list(NT, p["box"], p["Rc"] * p["cutoff"], p["Rn"] * p["cutoff"], p["skinrad"],
                          np.ascontiguousarray(s["x"][:NT]), np.ascontiguousarray(s["y"][:NT]),
                          np.ascontiguousarray(s["z"][:NT]), s["xd"], s["yd"], s["zd"], on, ol, inn, il,
                          p["numdims"])
    return ol[: on[0]], il[: inn[0]]


def _hip():
    """The HIP runtime libmpcx.so is linked against (the caller's own device
    memory comes from it: a second runtime in the process, such as the one
    torch bundles, would not see the GPU)."""
    lib()
    with open("/proc/self/maps") as f:
        paths = sorted({ln.split()[-1] for ln in f if "libamdhip64" in ln and "/torch/" not in ln})
    assert paths, "libamdhip64 not loaded"
    return C.CDLL(paths[0])


class _Dev:
    """A device buffer the test (the caller) allocates, filled from a host array."""

    def __init__(self, hip, a):
        self.hip, self.n = hip, a.nbytes
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), C.c_size_t(self.n)) == 0
        assert hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(self.n), 1) == 0   # to the device

    def read(self, dtype):
        out = np.empty(self.n // np.dtype(dtype).itemsize, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(self.n), 2) == 0
        return out

    def free(self):
        assert self.hip.hipFree(self.p) == 0


def _no_alloc(p, s, NT, pre):
    hip = _hip()
    xyz = [_Dev(hip, np.ascontiguousarray(s[k][:NT])) for k in ("x", "y", "z")]
    cap = NT + pre
    ol, il = _Dev(hip, np.full(cap, -7, np.int64)), _Dev(hip, np.full(cap, -7, np.int64))
    on, inn = np.array([pre], np.int64), np.array([pre], np.int64)
    try:
        D.inout_particle_list_noalloc(NT, p["box"], p["Rc"] * p["cutoff"], p["Rn"] * p["cutoff"], p["skinrad"],
                                      *[b.p.value for b in xyz], s["xd"], s["yd"], s["zd"], on, ol.p.value, cap,
                                      inn, il.p.value, cap, p["numdims"])
        return ol.read(np.int64)[: on[0]], il.read(np.int64)[: inn[0]]
    finally:
        for b in (*xyz, ol, il):
            b.free()


@pytest.mark.parametrize("numdims", [1, 2])
def test_no_alloc_lists_equal_the_allocating_entry_point(numdims):
    p, s = _system(numdims)
    NT = int(p["NT"])
    for n, pre in ((NT, 0), (NT // 3 + 5, 3)):   # the second call smaller, after 3 earlier entries
        ref_out, ref_in = _allocating(p, s, n, pre)
        out, inn = _no_alloc(p, s, n, pre)
        assert ref_in.size > 0 and ref_out.size > 0 and ref_in.size + ref_out.size == n + 2 * pre
        assert np.array_equal(out, ref_out) and np.array_equal(inn, ref_in)


def test_no_alloc_rejects_sizes_beyond_its_limits():
    p, s = _system(1)
    z = np.zeros(1, np.int64)
    for NT, nd, what in ((2 ** 26 + 1, 1, "67,108,864"), (10, 65, "64 dimers")):
        with pytest.raises(MpcxError, match=what):
            D.inout_particle_list_noalloc(NT, p["box"], 1.0, 1.0, 0.3, 0, 0, 0, np.zeros(2 * nd), np.zeros(2 * nd),
                                          np.zeros(2 * nd), z.copy(), 0, 0, z.copy(), 0, 0, nd)
