"""Throughput of the MD + MPC collision interval on one MI355X.

Workload (default): the plain FP64 collective_code system of the headline
configuration -- a 100x100x100-cell box with walls in z, 10 solvent per cell
(9,996,984 FP64 particles), one sphere dimer (Rc 2, Rn 4, bond 6.5), kT 1/6,
tau 0.5 = 50 MD steps of 0.01, skin 0.3, Hilbert-tour re-sort every 10
collisions, B->A bulk reaction k1 = 1e-3.  One "step" is one collision
interval: 50 MD steps plus the collision.  Synthetic initial state; the
collision draws come from the device Philox stream (the throughput mode; the
parity mode with host Mersenne-Twister draws is measured in the tests).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]

Prints one JSON line.  value = particle-updates/s (NT * K / wall time of the K
timed intervals, state resident in HBM).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_UPDATE_FP64 = 100.0   # pos+vel read (48), pos+vel written (48), flag read (4)


def headline_params(L=100):
    from main_amd import systems as S
    return S.collective_params(Lx=float(L), Ly=float(L), Lz=float(L), numdims=1, rhoa=10.0, Rc=2.0, Rn=4.0,
                               Dcn=6.5, EaC=1.0, EbC=1.0, EsC=1.0, EaN=1.0, EbN=0.1, EsN=1.0, T=1.0 / 6.0,
                               skinrad=0.3, MDt=0.01, MPCt=0.5, sort_step=500, k1=1e-3, react_box_type=7)


def cpu_baseline(p, s):
    """The oracle (C restatement of the reference, one core) on a bounded
    sample of the same workload: one collision interval of a 40^3-cell box
    with the same physics; reported in the bench metric's unit."""
    from oracle import oracle as O
    from main_amd import systems as S
    import copy
    ps = S.collective_params(Lx=40.0, Ly=40.0, Lz=40.0, numdims=1, rhoa=10.0, Rc=2.0, Rn=4.0, Dcn=6.5,
                             EbN=0.1, T=1.0 / 6.0, skinrad=0.3, MDt=0.01, MPCt=0.5, sort_step=1000000, k1=1e-3)
    ss = S.collective_initial_state(ps, seed=11)
    ss["mt"], ss["mti"] = O.mtmod_seed(4357787)
    O.co_run(ss, 0, do_setup=1)
    ncoll = 3
    t0 = time.perf_counter()
    O.co_run(ss, ncoll * ps["MPCtstep"])
    dt = time.perf_counter() - t0
    return {"value": ps["NT"] * ncoll / dt, "unit": "particle-updates/s", "cores": 1, "kind": "port",
            "sample": f"oracle, {ncoll} collision intervals of a 40^3-cell box ({ps['NT']} particles), "
                      f"{dt:.1f} s on one host core"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is described in DESIGN.md "
                                   "and not launched", "n_gpus": args.gpus}))
        sys.exit(2)

    from main_amd import systems as S
    from main_amd.collective import CollectiveEngine
    from main_amd.tour import cell_tour

    p = headline_params(args.size)
    t0 = time.time()
    s = S.collective_initial_state(p, seed=4357787)
    tour = cell_tour(int(p["box"][0]), int(p["box"][1]), int(p["box"][2]))
    print(f"# init {p['NT']} particles in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    eng = CollectiveEngine(p, rng_mode=1, seed=4357787)
    eng.set_state(s, tour=tour)
    eng.setup(do_sort=True)
    eng.run(args.warmup)
    print(f"# warmup {args.warmup} intervals done", file=sys.stderr, flush=True)
    from main_amd._lib import device_synchronize
    device_synchronize()
    t0 = time.perf_counter()
    eng.run(args.steps)
    device_synchronize()
    dt = time.perf_counter() - t0
    run_ms, bulk_ms, near_ms = eng.timing()
    print("# near engine per launch:", eng.near_profile(), file=sys.stderr, flush=True)
    st = eng.get_state()
    NT = p["NT"]
    value = NT * args.steps / dt
    ms_step = dt * 1000.0 / args.steps
    # roofline of the collision pipeline over all particles (bulk kernels of one interval)
    achieved = BYTES_PER_UPDATE_FP64 * NT / (bulk_ms * 1e-3) / 1e9
    out = {
        "metric": "particle-updates/s (MD+MPC)",
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "collective_code plain FP64, 100^3 cells with z walls, 10/cell "
                               f"({NT} solvent), 1 dimer, skin 0.3, tau 0.5 = 50 x 0.01, sort every 10 collisions",
                   "rng": "device Philox4x32-10", "md_steps_per_collision": p["MPCtstep"],
                   "particle_md_updates_per_s": value * p["MPCtstep"], "cstep": st["cstep"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "collision pipeline over all particles (co_bulk..co_rotate), 100 B/particle",
                     "bulk_ms": bulk_ms, "near_engine_ms": near_ms},
    }
    eng.close()
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(p, s)
        except Exception as ex:  # the oracle library is absent
            out["cpu_baseline"] = {"value": None, "error": str(ex)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
