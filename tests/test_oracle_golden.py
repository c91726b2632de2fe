"""The oracle against the reference's own results and generators.

reference_runs.json holds, for seeded small systems, the SHA-256 of every
array the reference's Fortran modules (compiled from /root/reference by
oracle/Makefile, one thread) produced; the oracle must reproduce each hash
bit for bit.  When oracle/_ref is built in this container the comparison is
also made live, array by array.
"""
import hashlib

import numpy as np
import pytest

from oracle import oracle as O
from tests.golden.make_golden import (COLLECTIVE_CASES, SELKOV_CASES, run_collective, run_selkov)


def h(a):
    return hashlib.sha256(np.ascontiguousarray(np.asarray(a)).tobytes()).hexdigest()


@pytest.mark.parametrize("name", sorted(COLLECTIVE_CASES))
def test_collective_oracle_matches_reference_hashes(golden, name):
    rec = golden["cases"][name]
    s = run_collective(COLLECTIVE_CASES[name], use_ref=False)
    bad = [k for k, v in rec["arrays"].items() if h(s[k]) != v]
    assert not bad, f"arrays differing from the reference: {bad}"
    for k, v in rec["scalars"].items():
        assert s[k] == v, k


@pytest.mark.parametrize("name", sorted(SELKOV_CASES))
def test_selkov_oracle_matches_reference_hashes(golden, name):
    rec = golden["cases"][name]
    s = run_selkov(SELKOV_CASES[name], use_ref=False)
    bad = [k for k, v in rec["arrays"].items() if h(s[k]) != v]
    assert not bad, f"arrays differing from the reference: {bad}"
    assert s["NT"] == rec["scalars"]["NT"]
    assert s["tpotential"] == rec["scalars"]["tpotential"]


@pytest.mark.skipif(O.ref_lib("collective") is None, reason="oracle/_ref not built (needs /root/reference)")
def test_collective_live_against_reference():
    case = COLLECTIVE_CASES["co_two_dimers_sort_wall"]
    a = run_collective(case, use_ref=False)
    b = run_collective(case, use_ref=True)
    for k in O.CO_ARRAYS + O.CO_MPC + ["mt"]:
        assert h(a[k]) == h(b[k]), k


@pytest.mark.skipif(O.ref_lib("selkov") is None, reason="oracle/_ref not built (needs /root/reference)")
def test_selkov_live_against_reference():
    case = SELKOV_CASES["sk_reactive"]
    a = run_selkov(case, use_ref=False)
    b = run_selkov(case, use_ref=True)
    for k in O.SK_ARRAYS + ["dm"]:
        assert h(a[k]) == h(b[k]), k


def test_mt_streams_match_reference_generators(golden):
    mt, mti = O.mtmod_seed(4357787)
    d, _, _ = O.mtmod_draws(mt, mti, 1000)
    assert h(d) == golden["mt"]["mtmod_4357787_first1000_sha256"]
    assert h(O.mt95_draws(8154543, 1000)) == golden["mt"]["mt95_8154543_first1000_sha256"]
