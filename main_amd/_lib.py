"""Loader of the HIP engine library (main_amd/libmpcx.so).

The product path has no CPU fallback: if the library is missing or no GPU is
visible the calls below raise.  Build it with ``python -c "import
__graft_entry__ as g; g.build()"`` (or ``make -C main_amd``).
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpcx.so")
_lib = None
# objects holding library handles: batches close before the engines they run
_live = (weakref.WeakSet(), weakref.WeakSet())


def track(obj, first: bool = False):
    """Registers an object with a close() method to be closed at interpreter
    exit, before the HIP runtime tears down (first: before the others)."""
    _live[0 if first else 1].add(obj)


@atexit.register
def _close_all():
    for group in _live:
        for obj in list(group):
            try:
                obj.close()
            except Exception:
                pass
    if _lib is not None:
        _lib.mpcx_shutdown()

P = C.c_void_p


def ptr(a):
    if a is None:
        return None
    if not isinstance(a, np.ndarray):
        raise TypeError("expected a numpy array")
    if not (a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]):
        raise ValueError("array must be contiguous")
    return a.ctypes.data_as(P)


class MpcxError(RuntimeError):
    pass


def lib():
    """The loaded libmpcx.so; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MpcxError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        _lib = C.CDLL(LIB_PATH)
        _lib.mpcx_last_error.restype = C.c_char_p
    return _lib


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().mpcx_last_error().decode(errors="replace")
        raise MpcxError(f"{what}: {msg} (code {rc})")


def device_count() -> int:
    return int(lib().mpcx_device_count())


def device_synchronize():
    check(lib().mpcx_device_synchronize(), "mpcx_device_synchronize")


def require_gpu():
    if device_count() < 1:
        raise MpcxError("no HIP device visible: the mpcx engine has no CPU path")


def philox_uniform2(seed: int, stream: int, step: int, idx0: int, n: int) -> np.ndarray:
    """(n, 2) uniforms of the device Philox stream the engines draw from in
    throughput mode: row i is the pair of counter index idx0 + i."""
    out = np.empty((n, 2), np.float64)
    check(lib().mpcx_philox_uniform2(C.c_uint64(seed), C.c_uint32(stream), C.c_uint64(step), C.c_uint64(idx0),
                                     C.c_int64(n), ptr(out)), "philox_uniform2")
    return out


# ---------------------------------------------------------------- host streams
class MtmodStream:
    """grnd() stream of collective_code/mtmod.f90 (state owned by the host)."""

    def __init__(self, seed: int | None = None, mt=None, mti=None):
        self.mt = np.zeros(624, np.uint32)
        self.mti = C.c_int32(0)
        if mt is not None:
            self.mt[:] = mt
            self.mti.value = int(mti)
        else:
            check(lib().mpcx_mtmod_seed(C.c_int32(seed), ptr(self.mt), C.byref(self.mti)), "mtmod_seed")

    def draws(self, n: int) -> np.ndarray:
        out = np.empty(n, np.float64)
        check(lib().mpcx_mtmod_fill(ptr(self.mt), C.byref(self.mti), ptr(out), C.c_int64(n)), "mtmod_fill")
        return out


class Mt95Stream:
    """genrand_real1 stream of selkov_code/mt95.f90."""

    def __init__(self, seed: int):
        self.val = np.zeros(624, np.uint32)
        self.cnt = C.c_int32(0)
        check(lib().mpcx_mt95_seed(C.c_int32(seed), ptr(self.val), C.byref(self.cnt)), "mt95_seed")

    def draws(self, n: int) -> np.ndarray:
        out = np.empty(n, np.float64)
        check(lib().mpcx_mt95_fill(ptr(self.val), C.byref(self.cnt), ptr(out), C.c_int64(n)), "mt95_fill")
        return out
