"""Writes the golden fixtures of tests/golden/ from the reference's own code.

Run in a container where /root/reference is mounted and oracle/_ref is built
(`make -C oracle ref`).  The reference's Fortran modules (collective_code and
selkov_code, single-thread) advance seeded systems built by
main_amd.systems; each fixture stores the recipe (parameters, seeds, step
counts) and the SHA-256 of every output array plus a few sample values.  The
tests rebuild the same inputs, run the oracle and compare the hashes, so the
fixtures pin the oracle to the reference without shipping large arrays.

Also records the first draws of the two Mersenne-Twister flavours, and a hash
of collective_code/hilbert.txt (the reference's only data vector).
"""
from __future__ import annotations

import copy
import hashlib
import json
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from main_amd import systems as S  # noqa: E402
from main_amd.tour import tour_coords  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"


def h(a) -> str:
    a = np.ascontiguousarray(np.asarray(a))
    return hashlib.sha256(a.tobytes()).hexdigest()


# the cases the tests replay (kept small: seconds on one core)
COLLECTIVE_CASES = {
    "co_one_dimer": dict(params=dict(Lx=16.0, Ly=16.0, Lz=20.0, rhoa=5.0, Rc=1.5, Rn=2.0, Dcn=4.0, MPCt=0.1,
                                     skinrad=0.3, sort_step=1000000, k1=0.05),
                         seed=3, mt_seed=4357787, steps=30, flags_mix=False, react_type=7, sort=False),
    "co_two_dimers_sort_wall": dict(params=dict(Lx=20.0, Ly=20.0, Lz=24.0, numdims=2, rhoa=4.0, rhob=1.0, Rc=1.5,
                                                Rn=2.0, Dcn=4.0, MPCt=0.1, skinrad=0.5, sort_step=20, k1=0.05,
                                                react_box_type=2, EbN=0.1),
                                    seed=11, mt_seed=4357787, steps=60, flags_mix=True, react_type=2, sort=True),
    "co_six_dimers_sort": dict(params=dict(Lx=28.0, Ly=28.0, Lz=16.0, numdims=6, rhoa=4.0, rhob=0.5, Rc=1.5,
                                           Rn=2.0, Dcn=4.0, MPCt=0.1, skinrad=0.3, sort_step=20, k1=0.05, EbN=0.1),
                               seed=21, mt_seed=4357787, steps=50, flags_mix=True, react_type=7, sort=True),
    # restarts (restart_mod.f95:416-495) in the middle of the run: lists and
    # forces evaluated again from the positions, with and without the sort
    "co_one_dimer_restart": dict(params=dict(Lx=16.0, Ly=16.0, Lz=20.0, rhoa=5.0, Rc=1.5, Rn=2.0, Dcn=4.0, MPCt=0.1,
                                             skinrad=0.3, sort_step=1000000, k1=0.05),
                                 seed=3, mt_seed=4357787, steps=40, restart_at=20, flags_mix=False, react_type=7,
                                 sort=False),
    "co_two_dimers_restart_sort": dict(params=dict(Lx=20.0, Ly=20.0, Lz=24.0, numdims=2, rhoa=4.0, rhob=1.0, Rc=1.5,
                                                   Rn=2.0, Dcn=4.0, MPCt=0.1, skinrad=0.5, sort_step=20, k1=0.05,
                                                   react_box_type=2, EbN=0.1),
                                       seed=11, mt_seed=4357787, steps=60, restart_at=40, flags_mix=True,
                                       react_type=2, sort=True),
    "co_six_dimers_restart": dict(params=dict(Lx=28.0, Ly=28.0, Lz=16.0, numdims=6, rhoa=4.0, rhob=0.5, Rc=1.5,
                                              Rn=2.0, Dcn=4.0, MPCt=0.1, skinrad=0.3, sort_step=20, k1=0.05,
                                              EbN=0.1),
                                  seed=21, mt_seed=4357787, steps=50, restart_at=30, flags_mix=True, react_type=7,
                                  sort=True),
}
SELKOV_CASES = {
    "sk_reactive": dict(params=dict(L=40.0, rho=2.0, Rc=2.0, Rn=4.0, Dcn=7.0, MDt=0.005, MPCt=0.1, NG=20000,
                                    k=(0.05, 0.01, 0.0004, 0.0004, 0.01, 0.01)),
                        seed=3, mt_seed=8154543, cstep0=1, steps=45, flags_mix=True),
}


def collective_inputs(case):
    p = S.collective_params(**case["params"])
    s = S.collective_initial_state(p, seed=case["seed"])
    if case["flags_mix"]:
        s["flag"][:] = np.random.default_rng(case["seed"] + 1).integers(0, 3, p["NT"]).astype(np.int32)
    rb = np.zeros(int(p["box"][0]) * int(p["box"][1]), np.int32)
    O.oracle_lib().or_co_reactive_grid_setup(O._ptr(p["box"]), O._ptr(rb), case["react_type"])
    s["react_box"] = rb
    s["mt"], s["mti"] = O.mtmod_seed(case["mt_seed"])
    return p, s


def selkov_inputs(case):
    p = S.selkov_params(**case["params"])
    s = S.selkov_initial_state(p, seed=case["seed"])
    if case["flags_mix"]:
        s["flag"][:] = np.random.default_rng(case["seed"] + 6).choice([0, 1, 3], p["NT"]).astype(np.int64)
    return p, s


def run_collective(case, use_ref):
    p, s = collective_inputs(case)
    L = [int(v) for v in p["box"]]
    tour = None
    if case["sort"]:
        c = tour_coords(*L).astype(np.int64)
        tour = c[:, 0] + L[0] * c[:, 1] + L[0] * L[1] * c[:, 2] + 1
        if use_ref:
            d = tempfile.mkdtemp()
            np.savetxt(os.path.join(d, "hilbert.txt"), c, fmt="%d", delimiter="\t")
            os.chdir(d)
    O.co_run(s, 0, do_setup=2 if case["sort"] else 1, use_ref=use_ref, tour=tour)
    ra = case.get("restart_at")
    if ra:
        O.co_run(s, ra, use_ref=use_ref, tour=tour)
        for k in ("fxd", "fyd", "fzd", "fpx", "fpy", "fpz", "pot"):   # a restart frame carries no forces
            s[k][...] = 0.0
        O.co_run(s, 0, do_setup=4 if (case["sort"] and ra % int(p["sort_step"]) == 0) else 3, use_ref=use_ref,
                 tour=tour)
        O.co_run(s, case["steps"] - ra, use_ref=use_ref, tour=tour)
    else:
        O.co_run(s, case["steps"], use_ref=use_ref, tour=tour)
    os.chdir(ROOT)
    return s


def run_selkov(case, use_ref):
    p, s = selkov_inputs(case)
    O.sk_run(s, case["cstep0"], case["steps"], seed=case["mt_seed"], use_ref=use_ref)
    return s


def main():
    if O.ref_lib("collective") is None or O.ref_lib("selkov") is None:
        sys.exit("oracle/_ref is not built: make -C oracle ref")
    out = {"cases": {}, "mt": {}, "hilbert": {}}
    for name, case in COLLECTIVE_CASES.items():
        ref = run_collective(case, True)
        ora = run_collective(case, False)
        rec = {"kind": "collective", "case": case, "arrays": {}, "scalars": {}}
        for k in O.CO_ARRAYS + O.CO_MPC + ["mt"]:
            assert h(ref[k]) == h(ora[k]), f"{name}: oracle differs from the reference in {k}"
            rec["arrays"][k] = h(ref[k])
        for k in ("cstep", "time_count", "time_penetrate", "out_num", "in_num", "mti"):
            rec["scalars"][k] = int(ref[k])
        rec["scalars"]["tpotential"] = float(ref["tpotential"])
        rec["samples"] = {k: np.asarray(ref[k]).ravel()[:4].tolist() for k in ("x", "vx", "xd", "vxd")}
        out["cases"][name] = rec
        print(name, "ok")
    for name, case in SELKOV_CASES.items():
        ref = run_selkov(case, True)
        ora = run_selkov(case, False)
        rec = {"kind": "selkov", "case": case, "arrays": {}, "scalars": {}}
        for k in O.SK_ARRAYS + ["dm"]:
            assert h(ref[k]) == h(ora[k]), f"{name}: oracle differs from the reference in {k}"
            rec["arrays"][k] = h(ref[k])
        rec["scalars"]["NT"] = int(ref["NT"])
        rec["scalars"]["tpotential"] = float(ref["tpotential"])
        out["cases"][name] = rec
        print(name, "ok")
    # RNG first draws from the reference's own generators
    lib = O.ref_lib("collective")
    mt = np.zeros(624, np.int32)
    mti = np.zeros(1, np.int32)
    lib.ref_mtmod_seed(O.C.c_int32(4357787), O._ptr(mt), O._ptr(mti))
    d = np.zeros(1000)
    lib.ref_mtmod_fill(O._ptr(mt), O._ptr(mti), O.C.c_int64(1000), O._ptr(d))
    out["mt"]["mtmod_4357787_first1000_sha256"] = h(d)
    out["mt"]["mtmod_4357787_first4"] = d[:4].tolist()
    d = np.zeros(1000)
    O.ref_lib("selkov").ref_mt95_fill(O.C.c_int32(8154543), O.C.c_int64(1000), O._ptr(d))
    out["mt"]["mt95_8154543_first1000_sha256"] = h(d)
    out["mt"]["mt95_8154543_first4"] = d[:4].tolist()
    hil = np.loadtxt(os.path.join(REF, "collective_code", "hilbert.txt"), dtype=np.int32)
    out["hilbert"] = {"shape": list(hil.shape), "sha256_int32": h(hil), "first": hil[:3].tolist(),
                      "last": hil[-3:].tolist()}
    with open(os.path.join(OUT, "reference_runs.json"), "w") as f:
        json.dump(out, f, indent=1, default=lambda o: list(o) if isinstance(o, tuple) else o)
    print("wrote", os.path.join(OUT, "reference_runs.json"))


if __name__ == "__main__":
    main()
