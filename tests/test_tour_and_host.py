"""Cell tours, derived parameters and the C ABI of libmpcx (no GPU needed)."""
import ctypes as C
import hashlib
import os
import re

import numpy as np
import pytest

from main_amd import systems as S
from main_amd.tour import cell_tour, hilbert_curve, reference_tour_coords_64x64x32, tour_coords

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_tour_reproduces_reference_hilbert_txt(golden):
    t = reference_tour_coords_64x64x32().astype(np.int32)
    g = golden["hilbert"]
    assert list(t.shape) == g["shape"]
    assert hashlib.sha256(np.ascontiguousarray(t).tobytes()).hexdigest() == g["sha256_int32"]
    assert t[:3].tolist() == g["first"] and t[-3:].tolist() == g["last"]


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_hilbert_curve_is_unit_step_permutation(k):
    c = hilbert_curve(k)
    n = 1 << k
    idx = c[:, 0] + n * c[:, 1] + n * n * c[:, 2]
    assert np.array_equal(np.sort(idx), np.arange(n ** 3))
    steps = np.abs(np.diff(c, axis=0)).sum(1)
    assert np.all(steps == 1)


@pytest.mark.parametrize("L", [(30, 30, 30), (100, 100, 100), (200, 200, 10), (23, 37, 31)])
def test_cell_tour_is_permutation(L):
    t = cell_tour(*L)
    assert np.array_equal(np.sort(t), np.arange(1, L[0] * L[1] * L[2] + 1))


def test_config_sizes_match_reference_counts():
    # main_pattern.f95:66-73 with Rc=2, Rn=4, rho=10 (the counts quoted for the five configurations)
    for L, NT in (((30, 30, 30), 266984), ((100, 100, 100), 9996984), ((50, 50, 50), 1246984)):
        p = S.collective_params(Lx=float(L[0]), Ly=float(L[1]), Lz=float(L[2]), rhoa=10.0)
        assert p["NT"] == NT
    p = S.collective_params(Lx=100.0, Ly=100.0, Lz=100.0)
    assert p["mpcnumcell"] == 1010000 and p["rxnnumcell"] == 1000000 and p["MPCtstep"] == 50


def _header_functions():
    with open(os.path.join(ROOT, "include", "mpcx.h")) as f:
        src = f.read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mpcx_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from main_amd._lib import LIB_PATH
    assert os.path.exists(LIB_PATH), "libmpcx.so not built"
    lib = C.CDLL(LIB_PATH)
    missing = [f for f in _header_functions() if not hasattr(lib, f)]
    assert not missing, f"declared in include/mpcx.h but not exported: {missing}"


def test_product_mt_streams_match_oracle():
    from main_amd._lib import MtmodStream, Mt95Stream
    from oracle import oracle as O
    mt, mti = O.mtmod_seed(4357787)
    ref, _, _ = O.mtmod_draws(mt, mti, 5000)
    assert np.array_equal(MtmodStream(4357787).draws(5000), ref)
    assert np.array_equal(Mt95Stream(8154543).draws(5000), O.mt95_draws(8154543, 5000))


def test_co_config_layout():
    from main_amd.collective import CoConfig
    # 6 int64 + 4 int32 + uint64 + 3 doubles + 30 doubles + 101 doubles + 1 double
    assert C.sizeof(CoConfig) == 6 * 8 + 4 * 4 + 8 + 3 * 8 + 30 * 8 + 101 * 8 + 8


def test_reaction_table_matches_reference_formula():
    from main_amd.collective import reaction_table
    t = reaction_table(1e-3, 0.5)
    assert t[0] == 0.0
    assert t[10] == 1.0 - np.exp(-(1e-3 * 10.0 * 0.5))


def test_exit_hook_closes_batches_before_engines(monkeypatch):
    """At interpreter exit every object holding a library handle is closed,
    batches before the engines they run, before the library's own
    mpcx_shutdown and the HIP runtime's teardown."""
    import main_amd._lib as L
    order = []

    class Obj:
        def __init__(self, name):
            self.name = name

        def close(self):
            order.append(self.name)

    objs = [Obj("engine"), Obj("batch")]
    L.track(objs[0])
    L.track(objs[1], first=True)
    monkeypatch.setattr(L, "_lib", None)
    L._close_all()
    assert order == ["batch", "engine"]
