"""ctypes front end of the CPU checker -- TEST INFRASTRUCTURE ONLY.

Loads oracle/liboracle.so (the C restatement of the reference algorithm) and,
when it was built in this container, oracle/_ref/libref_*.so (the reference's
own Fortran modules behind a bind(C) harness).  Both expose the same flat
driver signatures, so a test can run one system through either and compare
every array.  Only tests/, the smoke entry and bench.py's CPU-baseline leg
import this module; the product path never does.
"""
from __future__ import annotations

import ctypes as C
import os
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MAXS_CO = 100   # MaxA/MaxB/MaxS, collective_code/global_mod.f95:114-116
MAXS_SK = 200   # MaxF/MaxG/MaxI, selkov_code/Global_SelkovDimer.f90:129-131

_P = C.c_void_p


def _ptr(a: np.ndarray):
    assert a.flags["F_CONTIGUOUS"] or a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_P)


def _load(path):
    if not os.path.exists(path):
        return None
    return C.CDLL(path)


# ------------------------------------------------------------------ Philox
# Philox4x32-10 of Salmon, Moraes, Dror and Shaw ("Parallel random numbers: as
# easy as 1, 2, 3", SC'11), restated from the paper in numpy: ten rounds of
# two 32x32->64-bit products on counter words 0 and 2, the key bumped by the
# Weyl constants after each round.  The engines draw from it in throughput
# mode (main_amd/csrc/mpcx_device.h: philox4x32_10, philox_uniform2); the
# reference itself uses Mersenne Twister and has no counterpart.
PHILOX_M = (0xD2511F53, 0xCD9E8D57)
PHILOX_W = (0x9E3779B9, 0xBB67AE85)


def philox4x32_10(ctr, key):
    """ctr: (n, 4) uint32 counters; key: two uint32 words -> (n, 4) uint32."""
    c = np.array(ctr, dtype=np.uint64).reshape(-1, 4)
    k0, k1 = np.uint64(key[0]), np.uint64(key[1])
    m0, m1, mask = np.uint64(PHILOX_M[0]), np.uint64(PHILOX_M[1]), np.uint64(0xFFFFFFFF)
    for _ in range(10):
        p0 = m0 * c[:, 0]
        p1 = m1 * c[:, 2]
        n = np.empty_like(c)
        n[:, 0] = (p1 >> np.uint64(32)) ^ c[:, 1] ^ k0
        n[:, 1] = p1 & mask
        n[:, 2] = (p0 >> np.uint64(32)) ^ c[:, 3] ^ k1
        n[:, 3] = p0 & mask
        c = n
        k0 = (k0 + np.uint64(PHILOX_W[0])) & mask
        k1 = (k1 + np.uint64(PHILOX_W[1])) & mask
    return c.astype(np.uint32)


def philox_uniform2(seed: int, stream: int, step: int, idx):
    """The pair of 53-bit uniforms of counter (idx, step, stream) under key
    seed, as the device draws them: (n, 2) float64."""
    idx = np.asarray(idx, dtype=np.uint64).reshape(-1)
    ctr = np.empty((idx.size, 4), np.uint64)
    ctr[:, 0] = idx & np.uint64(0xFFFFFFFF)
    ctr[:, 1] = idx >> np.uint64(32)
    ctr[:, 2] = np.uint64(step & 0xFFFFFFFF)
    ctr[:, 3] = np.uint64(((step >> 32) & 0xFFFF) | (stream << 16))
    r = philox4x32_10(ctr, (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)).astype(np.uint64)
    a = (r[:, 0] << np.uint64(21)) ^ (r[:, 1] >> np.uint64(11))
    b = (r[:, 2] << np.uint64(21)) ^ (r[:, 3] >> np.uint64(11))
    return np.stack([a, b], axis=1).astype(np.float64) * 2.0 ** -53


def oracle_lib():
    lib = _load(os.path.join(HERE, "liboracle.so"))
    if lib is None:
        raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle oracle`)")
    lib.or_mtmod_grnd.restype = C.c_double
    lib.or_mt95_real1.restype = C.c_double
    lib.or_mt95_int32.restype = C.c_uint32
    return lib


def ref_lib(kind: str):
    """The flang-built reference ('collective', 'selkov', or 'collective_omp', the
    collective modules built with their OpenMP directives on), or None if absent."""
    return _load(os.path.join(HERE, "_ref", f"libref_{kind}.so"))


# ----------------------------------------------------------------- RNG helpers
class MtmodState(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("mti", C.c_int32)]


class Mt95State(C.Structure):
    _fields_ = [("val", C.c_uint32 * 624), ("cnt", C.c_int32)]


def mtmod_seed(seed: int):
    """(mt[624] uint32, mti) after sgrnd(seed) (mtmod.f90:72-85)."""
    lib = oracle_lib()
    s = MtmodState()
    lib.or_mtmod_sgrnd(C.byref(s), C.c_int32(seed))
    return np.frombuffer(bytes(s.mt), dtype=np.uint32).copy(), int(s.mti)


def mtmod_draws(mt: np.ndarray, mti: int, n: int):
    lib = oracle_lib()
    s = MtmodState()
    C.memmove(s.mt, mt.ctypes.data, 624 * 4)
    s.mti = mti
    out = np.empty(n, np.float64)
    lib.or_mtmod_fill(C.byref(s), _ptr(out), C.c_int64(n))
    return out, np.frombuffer(bytes(s.mt), dtype=np.uint32).copy(), int(s.mti)


def mt95_draws(seed: int, n: int):
    lib = oracle_lib()
    s = Mt95State()
    lib.or_mt95_init(C.byref(s), C.c_int32(seed))
    out = np.empty(n, np.float64)
    lib.or_mt95_fill(C.byref(s), _ptr(out), C.c_int64(n))
    return out


def mt95_int32(seed: int, n: int):
    lib = oracle_lib()
    s = Mt95State()
    lib.or_mt95_init(C.byref(s), C.c_int32(seed))
    return np.array([lib.or_mt95_int32(C.byref(s)) for _ in range(n)], dtype=np.uint32)


# ----------------------------------------------------------- collective driver
CO_ARRAYS = ["flag", "x", "y", "z", "vx", "vy", "vz", "xd", "yd", "zd", "vxd", "vyd", "vzd",
             "fxd", "fyd", "fzd", "fpx", "fpy", "fpz", "pot", "nb", "nbnum", "msd", "ol", "il"]
CO_MPC = ["tA", "tB", "tS", "sA", "sB", "sS", "vxcm", "vycm", "vzcm", "tcs", "ncn", "uvx", "uvy", "uvz"]


def co_blocks(sysd: dict):
    """Pack the scalar blocks (layout documented in ref_collective_harness.f90)."""
    ip = np.array([sysd["NT"], sysd["numdims"], sysd["MPCtstep"], sysd["sort_step"], sysd["react_dir"],
                   sysd["cstep"], sysd["time_count"], sysd["time_penetrate"], sysd["vol"], sysd["max_nb"],
                   sysd["out_num"], sysd["in_num"], sysd["mpcnumcell"], sysd["rxnnumcell"]], dtype=np.int64)
    rp = np.zeros(31, np.float64)
    rp[0:3] = sysd["box"]
    keys = ["MDt", "Ma", "Mb", "Ms", "Mc", "Mn", "Dcn", "cutoff", "Rc", "Rn", "EaC", "EbC", "EsC", "EaN",
            "EbN", "EsN", "Rw", "Ew", "Rm_delta", "Em", "RPab", "RPba", "skinrad", "rotang", "a0", "MPCt", "k1"]
    for i, k in enumerate(keys):
        rp[3 + i] = sysd[k]
    return ip, rp


def co_alloc_work(sysd: dict):
    """Arrays of the reference's shapes that the run writes (lists, census, cells)."""
    NT, nd, mnb, vol, mc = sysd["NT"], sysd["numdims"], sysd["max_nb"], sysd["vol"], sysd["mpcnumcell"]
    w = {}
    for k in ("fpx", "fpy", "fpz", "pot"):
        w.setdefault(k, np.zeros((nd, 2, mnb), order="F"))
    w.setdefault("nb", np.zeros((nd, 2, mnb), np.int64, order="F"))
    w.setdefault("nbnum", np.zeros((nd, 2), np.int64, order="F"))
    w.setdefault("msd", np.zeros(NT))
    w.setdefault("ol", np.zeros(NT, np.int64))
    w.setdefault("il", np.zeros(2 * nd * mnb, np.int64))
    for k in ("tA", "tB", "tS"):
        w[k] = np.zeros(vol, np.int64)
    for k in ("sA", "sB", "sS"):
        w[k] = np.zeros((vol, MAXS_CO), np.int64, order="F")
    for k in ("vxcm", "vycm", "vzcm", "tcs", "uvx", "uvy", "uvz"):
        w[k] = np.zeros(mc)
    w["ncn"] = np.zeros(NT, np.int64)
    return w


def co_run(sysd: dict, nsteps: int, do_setup: int = 0, use_ref: bool = False, tour=None):
    """Advance a collective system in place by nsteps MD steps.

    sysd holds parameters, state arrays (Fortran layouts), the mtmod state
    ("mt", "mti") and the scalar loop state; it is updated in place.  With
    use_ref the flang-built reference modules run (sort_hilbert then reads
    hilbert.txt from the working directory); otherwise the C restatement runs
    with the tour given as 1-based cell numbers.
    """
    threaded = use_ref == "omp"
    lib = ref_lib("collective_omp" if threaded else "collective") if use_ref else oracle_lib()
    if lib is None:
        raise RuntimeError("reference library not built")
    for k, v in co_alloc_work(sysd).items():
        if k not in sysd or sysd[k] is None or np.shape(sysd[k]) != np.shape(v):
            sysd[k] = v
    ip, rp = co_blocks(sysd)
    if threaded:   # the OpenMP build: one stream per thread, mt(624, nthreads) column-major
        mt = np.ascontiguousarray(sysd["mt_threads"], dtype=np.uint32)
        mti = np.ascontiguousarray(sysd["mti_threads"], dtype=np.int32)
        lib.ref_set_mt_threads(C.c_int32(mti.size))
        lib.omp_set_num_threads(C.c_int(mti.size))
    else:
        mt = np.ascontiguousarray(sysd["mt"], dtype=np.uint32)
        mti = np.array([sysd["mti"]], dtype=np.int32)
    rb = np.ascontiguousarray(sysd["react_box"], dtype=np.int32)
    args = [_ptr(ip), _ptr(rp), C.c_int64(nsteps), C.c_int32(do_setup)]
    args += [_ptr(sysd[k]) for k in CO_ARRAYS]
    args += [C.c_int64(rb.size), _ptr(rb)]
    args += [_ptr(sysd[k]) for k in CO_MPC]
    args += [_ptr(mt), _ptr(mti)]
    if use_ref:
        lib.ref_co_run(*args)
        rc = 0
    else:
        t = np.ascontiguousarray(tour if tour is not None else np.arange(1, sysd["vol"] + 1), dtype=np.int64)
        lib.or_co_run.restype = C.c_int
        rc = lib.or_co_run(*args, _ptr(t))
    if threaded:
        sysd["mt_threads"], sysd["mti_threads"] = mt, mti
    else:
        sysd["mt"], sysd["mti"] = mt, int(mti[0])
    (sysd["cstep"], sysd["time_count"], sysd["time_penetrate"]) = (int(ip[5]), int(ip[6]), int(ip[7]))
    sysd["out_num"], sysd["in_num"] = int(ip[10]), int(ip[11])
    sysd["tpotential"] = float(rp[30])
    if rc != 0:
        raise RuntimeError("oracle reported an out-of-range cell or member count")
    return sysd


# --------------------------------------------------------------- selkov driver
def sk_blocks(sysd: dict):
    ip = np.array([sysd["NT"], sysd["space"], sysd.get("sort_step", 10000)], dtype=np.int64)
    rp = np.zeros(29, np.float64)
    rp[0:3] = sysd["box"]
    keys = ["Mf", "Mg", "Mi", "Rc", "Rn", "Dcn", "EfC", "EgC", "EiC", "EfN", "EgN", "EiN", "RPfg", "RPgf",
            "A0", "MDt", "MPCt", "k1", "k2", "k3", "k4", "k5", "k6", "Mc", "Mn"]
    for i, k in enumerate(keys):
        rp[3 + i] = sysd[k]
    return ip, rp


SK_ARRAYS = ["flag", "x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz", "pfx", "pfy", "pfz"]


def sk_run(sysd: dict, cstep0: int, nsteps: int, seed: int, use_ref: bool = False):
    """Run Main_SelkovDimer.f90:125-149 for nsteps steps from cstep0 with a
    fresh mt95 stream seeded by `seed`; state arrays updated in place."""
    lib = ref_lib("selkov") if use_ref else oracle_lib()
    if lib is None:
        raise RuntimeError("reference library not built")
    ip, rp = sk_blocks(sysd)
    dm = np.asfortranarray(sysd["dm"], dtype=np.float64)
    args = [_ptr(ip), _ptr(rp), C.c_int32(seed), C.c_int64(cstep0), C.c_int64(nsteps)]
    args += [_ptr(sysd[k]) for k in SK_ARRAYS] + [_ptr(dm)]
    if use_ref:
        lib.ref_sk_run(*args)
        rc = 0
    else:
        lib.or_sk_run.restype = C.c_int
        rc = lib.or_sk_run(*args)
    sysd["dm"] = dm
    sysd["NT"] = int(ip[0])
    sysd["tpotential"] = float(rp[28])
    if rc != 0:
        raise RuntimeError("oracle reported an out-of-range cell or member count")
    return sysd
