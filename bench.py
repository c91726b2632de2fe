"""Throughput of the MD + MPC collision interval on one MI355X.

Workload (default): the plain FP64 collective_code system of the headline
configuration -- a 100x100x100-cell box with walls in z, 10 solvent per cell
(9,996,984 FP64 particles), one sphere dimer (Rc 2, Rn 4, bond 6.5), kT 1/6,
tau 0.5 = 50 MD steps of 0.01, skin 0.3, Hilbert-tour re-sort every 10
collisions, B->A bulk reaction k1 = 1e-3.  One "step" is one collision
interval: 50 MD steps plus the collision.  Synthetic initial state; the
collision draws come from the device Philox stream (the throughput mode; the
parity mode with host Mersenne-Twister draws is measured in the tests).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline] [--probe-copy]
    python bench.py --summarize-pmc FETCH_DIR WRITE_DIR    (rocprofv3 counter CSVs -> traffic)

Prints one JSON line.  value = particle-updates/s (NT * K / wall time of the K
timed intervals, state resident in HBM).
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROBE_BYTES = 256 << 20          # copy probe: reads and writes this many bytes per launch
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_100cube.txt")
TRAFFIC_SPLIT_FILE = os.path.join(ROOT, "profiles", "traffic_split_batched_256.txt")

# Algorithmic HBM bytes per particle of each interval kernel (FP64 state):
#  co_bulk   reads x,y,z,vx,vy,vz (48), flag (4), status (1); writes x,y,z (24),
#            collision-cell position (4), rank (2), distance byte (1)        = 84
#  co_rotate reads cell position (4), vx,vy,vz (24), distance byte (1); writes vx,vy,vz (24)
#            (+ the cell records, 64 B per cell, about 6.4 B per particle)  = 59.4
#  co_collide per particle: bucket entry (4) + vx,vy,vz gathered (24); per cell 64 B written
#  co_scatter reads cell (4) + rank (2) + offset (4), writes bucket entry (4)     = 14
#  co_far    the bytes of co_bulk for the far particles (all but the ~4,300 of the near
#            set at 100^3, counted as NT), run beside the near engine               = 84
ALGO_BYTES = {"co_bulk": 84.0, "co_far": 84.0, "co_rotate": 59.4, "co_collide": 34.4, "co_scatter": 14.0}


def headline_params(L=100):
    from main_amd import systems as S
    return S.collective_params(Lx=float(L), Ly=float(L), Lz=float(L), numdims=1, rhoa=10.0, Rc=2.0, Rn=4.0,
                               Dcn=6.5, EaC=1.0, EbC=1.0, EsC=1.0, EaN=1.0, EbN=0.1, EsN=1.0, T=1.0 / 6.0,
                               skinrad=0.3, MDt=0.01, MPCt=0.5, sort_step=500, k1=1e-3, react_box_type=7)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_leg(L, budget_s, threads):
    """Whole collision intervals of the L^3 system on the CPU, as many as fit
    the budget: the flang-built reference modules (threads > 1: their OpenMP
    build, one Mersenne-Twister stream per thread as the reference keeps
    them), else the oracle's C restatement."""
    from oracle import oracle as O
    from main_amd import systems as S
    kind = "omp" if threads > 1 else True
    use_ref = O.ref_lib("collective_omp" if threads > 1 else "collective") is not None
    if not use_ref and threads > 1:
        return None
    ps = headline_params(L)
    ps["sort_step"] = 1000000   # no re-sort inside the sample
    ss = S.collective_initial_state(ps, seed=11)
    if threads > 1:
        seeds = [O.mtmod_seed(4357787 + t) for t in range(threads)]
        ss["mt_threads"] = np.stack([m for m, _ in seeds]).astype(np.uint32)
        ss["mti_threads"] = np.array([i for _, i in seeds], np.int32)
    else:
        ss["mt"], ss["mti"] = O.mtmod_seed(4357787)
    mode = kind if use_ref else False
    O.co_run(ss, 0, do_setup=1, use_ref=mode)
    n, dt = 0, 0.0
    while dt < budget_s and n < 20:
        t0 = time.perf_counter()
        O.co_run(ss, ps["MPCtstep"], use_ref=mode)
        dt += time.perf_counter() - t0
        n += 1
    what = ("flang-built reference modules" + (f", OpenMP build, {threads} threads" if threads > 1 else ", one thread")
            if use_ref else "C restatement of the reference, one thread")
    return {"value": ps["NT"] * n / dt, "unit": "particle-updates/s", "cores": threads,
            "kind": "reference" if use_ref else "port",
            "sample": f"{what}: {n} collision intervals of the {L}^3 system ({ps['NT']} particles), {dt:.1f} s"}


def cpu_baseline(L=100, budget_s=15.0):
    """The reference's own Fortran modules on a bounded sample of the same
    workload, on this host's cores: at the core share (OMP_NUM_THREADS, else
    every core; the reference's jobs run OpenMP, collective_code/Run.sh:10)
    and at one thread.  Falls back to the oracle's C restatement on one thread
    when the reference build is absent."""
    ncpu = os.cpu_count() or 1
    threads = max(1, min(ncpu, int(os.environ.get("OMP_NUM_THREADS", ncpu) or ncpu)))
    one = _cpu_leg(L, budget_s, 1)
    many = _cpu_leg(L, budget_s, threads) if threads > 1 else None
    out = dict(many or one)
    out["cpu"] = cpu_model()
    out["host_cpus"] = ncpu
    if many:
        out["single_thread"] = one
        out["thread_speedup"] = many["value"] / one["value"]
    return out


def batched(R=256, steps=20, warmup=4, seed=4357787, tiles=(1024, 2048, 4096, 8192)):
    """The replica ensemble in FP64 storage, advanced together by one batch
    (HIP-graph replay, autotuned tile)."""
    from main_amd.batch import CollectiveBatch
    from main_amd._lib import device_synchronize
    t0 = time.time()
    engines = ensemble_engines(R, seed, False)
    print(f"# batched: {R} replicas built in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    batch = CollectiveBatch(engines, tile=2048)
    info = batch.info()
    times = batch.autotune(tiles=tiles, ncoll=10)
    batch.run(warmup)
    device_synchronize()
    t0 = time.perf_counter()
    batch.run(steps)
    device_synchronize()
    dt = time.perf_counter() - t0
    ms = dt * 1000.0 / steps
    out = {"replicas": R, "particles": info["particles"], "collision_cells": info["cells"],
           "ms_per_interval": ms, "particle_updates_per_s": info["particles"] * steps / dt,
           "replicas_per_hour_2000_collisions": R * 3600.0 / (2000 * ms * 1e-3),
           "tile": batch.tile, "autotune_ms_per_interval": times, "rng": "device Philox4x32-10",
           "graph": True, "dtype": "f64", "roofline_step": step_roofline(FP64_STEP_BYTES, info["particles"], ms)}
    batch.close()
    for e in engines:
        e.close()
    return out


MANY_DIMER = {
    # 200x200x10 slab, 100 dimers in the mid-plane, plain FP64 (the split-precision
    # configuration's geometry and physics)
    "slab_100_dimers_200x200x10": dict(Lx=200.0, Ly=200.0, Lz=10.0, numdims=100),
    # collective_code/input.txt: 64x64x32 box, 20 dimers
    "reference_input_20_dimers_64x64x32": dict(Lx=64.0, Ly=64.0, Lz=32.0, numdims=20),
}


def many_dimers(name, steps=20, warmup=5, seed=4357787):
    """One many-dimer system (device-generated initial solvent, headline
    physics, sort every 10 collisions), advanced by the global near engine."""
    from main_amd import systems as S
    from main_amd._lib import device_synchronize
    from main_amd.collective import CollectiveEngine
    from main_amd.tour import cell_tour
    p = S.collective_params(**MANY_DIMER[name], rhoa=10.0, Rc=2.0, Rn=4.0, Dcn=6.5, EbN=0.1, T=1.0 / 6.0,
                            skinrad=0.3, MDt=0.01, MPCt=0.5, sort_step=500, k1=1e-3)
    t0 = time.time()
    d = S.collective_dimers(p, seed=seed)
    eng = CollectiveEngine(p, rng_mode=1, seed=seed)
    eng.set_dimers(d, tour=cell_tour(*[int(v) for v in p["box"]]))
    eng.init_solvent(seed)
    eng.setup(do_sort=True)
    device_synchronize()
    t_init = time.time() - t0
    eng.run(warmup)
    device_synchronize()
    t0 = time.perf_counter()
    eng.run(steps)
    device_synchronize()
    dt = time.perf_counter() - t0
    kms = eng.kernel_ms()
    prof = eng.near_profile()
    eng.close()
    print(f"# {name}: workgroup-0 us per launch: total {prof['total_us']:.0f} rebuild {prof['rebuild_us']:.0f} "
          f"particles {prof['particle_us']:.0f} dimers {prof['serial_us']:.0f} rebuilds {prof['rebuilds']:.1f} "
          f"barrier waits {prof['near_num'] / 100.0 / max(prof['launches'], 1):.0f} "
          f"totals {prof['inside_num'] / 100.0 / max(prof['launches'], 1):.0f}", file=sys.stderr, flush=True)
    print(f"# {name}: {dt * 1000.0 / steps:.2f} ms/interval", file=sys.stderr, flush=True)
    return {"particles": p["NT"], "dimers": p["numdims"], "ms_per_interval": dt * 1000.0 / steps,
            "particle_updates_per_s": p["NT"] * steps / dt, "init_s": t_init, "kernel_ms": kms,
            "near_engine": "global (cooperative workgroups)", "dtype": "f64", "rng": "device Philox4x32-10"}


def selkov(steps=20, warmup=5, seed=8154543):
    """selkov_code at 50^3 cells, 10 per cell (1.25e6 particles), one dimer,
    the reference input's rates, tau 0.1 = 20 x 0.005, device Philox draws."""
    from main_amd import systems as S
    from main_amd._lib import device_synchronize
    from main_amd.selkov import SelkovEngine
    p = S.selkov_params(L=50.0, rho=10.0, NG=100000)
    s = S.selkov_initial_state(p, seed=seed)
    eng = SelkovEngine(p, rng_mode=1, seed=seed)
    eng.set_state(s, cstep=0)
    eng.run(warmup)
    device_synchronize()
    t0 = time.perf_counter()
    eng.run(steps)
    device_synchronize()
    dt = time.perf_counter() - t0
    kms = eng.kernel_ms()
    eng.close()
    print(f"# selkov 50^3: {dt * 1000.0 / steps:.3f} ms/interval", file=sys.stderr, flush=True)
    return {"particles": p["NT"], "md_steps_per_collision": p["MPCtstep"], "ms_per_interval": dt * 1000.0 / steps,
            "particle_updates_per_s": p["NT"] * steps / dt, "kernel_ms": kms, "dtype": "f64",
            "rng": "device Philox4x32-10"}


# Algorithmic bytes per particle of the split-precision interval kernels (29-byte state: packed
# int32 cell, 3 x FP32 offsets, 3 x FP32 velocities, int8 species):
#  sb_bulk   far particles: reads cell, offsets, velocities, species (29), status (1); writes cell and
#            offsets (16), collision cell (4), distance byte (1)                               = 51
#  sb_rotate reads collision cell (4), velocities (12), distance byte (1), the cell record (32 B per
#            cell, 3.2 per particle); writes velocities (12)                                  = 32.2
#  sb_cells_own  per collision cell: the owner's sums read (24), one accumulator copy read and cleared
#            (48), record written (32); per unshifted cell: B census and first key read and cleared (16);
#            per particle (10 per cell)                                                    = 12.0
# (the owner tiles' stores of their cells' sums, 2.4 B per particle, and the inbox entries between
# tiles are the binning's own traffic, not counted as algorithmic)
# The interval's figure is 57 B per particle-update: the split state read once (cell 4, offsets 12,
# velocities 12, species 1) and written once (the same without the species, written only on a change).
ALGO_BYTES_SPLIT = {"sb_bulk": 51.0, "sb_rotate": 32.2, "sb_cells_own": 12.0}
SPLIT_STEP_BYTES = 57.0
FP64_STEP_BYTES = 100.0


def split_headline(L=100, steps=60, warmup=20, seed=4357787):
    """The headline system in split-precision storage (same physics, device
    Philox): a one-system batch of the split interval."""
    from main_amd import systems as S
    from main_amd._lib import device_synchronize
    from main_amd.collective import CollectiveEngine
    from main_amd.tour import cell_tour
    p = headline_params(L)
    d = S.collective_dimers(p, seed=seed)
    eng = CollectiveEngine(p, rng_mode=1, seed=seed, split=True)
    eng.set_dimers(d, tour=cell_tour(L, L, L))
    eng.init_solvent(seed)
    eng.setup(do_sort=True)
    eng.run(warmup)
    device_synchronize()
    t0 = time.perf_counter()
    eng.run(steps)
    device_synchronize()
    dt = time.perf_counter() - t0
    kms = eng.kernel_ms()
    eng.close()
    NT = p["NT"]
    ms = dt * 1000.0 / steps
    print(f"# split {L}^3: {ms:.3f} ms/interval {kms}", file=sys.stderr, flush=True)
    return {"particles": NT, "ms_per_interval": ms, "particle_updates_per_s": NT * steps / dt,
            "kernel_ms": kms, "dtype": "f32 solvent (cell + offsets), f64 near set and dimer",
            "roofline_step": step_roofline(SPLIT_STEP_BYTES, NT, ms),
            "kernels": kernel_rooflines(kms, ALGO_BYTES_SPLIT, NT)}


def step_roofline(bytes_per_particle, particles, ms_per_interval):
    """The whole interval against the HBM roofline: algorithmic bytes of one
    particle-update times the particles, over the interval's wall time."""
    ach = bytes_per_particle * particles / (ms_per_interval * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "algorithmic_bytes_per_particle": bytes_per_particle}


def kernel_rooflines(kms, algo, particles):
    out = {}
    for k, b in algo.items():
        t = kms.get(k)
        if t:
            ach = b * particles / (t * 1e-3) / 1e9
            out[k] = {"launch_ms": t, "achieved": ach, "frac": ach / HBM_PEAK_GBS, "algorithmic_bytes_per_particle": b}
    return out


def ensemble_engines(R, seed, split):
    """The replica ensemble: R single-dimer systems with cubic boxes of 20-40
    cells (edges uniform), 10 solvent per cell, RPab uniform in [0.2, 1], EbN
    uniform in [0.1, 1], one Philox key per replica, the other physics as the
    headline system (re-sort every 10 collisions); device initial solvent."""
    from main_amd import systems as S
    from main_amd.collective import CollectiveEngine
    from main_amd.tour import cell_tour
    rng = np.random.default_rng(seed)
    engines, tours = [], {}
    for r in range(R):
        L = int(rng.integers(20, 41))
        p = S.collective_params(Lx=float(L), Ly=float(L), Lz=float(L), numdims=1, rhoa=10.0, Rc=2.0, Rn=4.0,
                                Dcn=6.5, EbN=float(rng.uniform(0.1, 1.0)), RPab=float(rng.uniform(0.2, 1.0)),
                                T=1.0 / 6.0, skinrad=0.3, MDt=0.01, MPCt=0.5, sort_step=500, k1=1e-3)
        if L not in tours:
            tours[L] = cell_tour(L, L, L)
        e = CollectiveEngine(p, rng_mode=1, seed=seed + r, split=split)
        e.set_dimers(S.collective_dimers(p, seed=seed + r), tour=tours[L])
        e.init_solvent(seed + r)
        e.setup(do_sort=True)
        engines.append(e)
    return engines


def split_batched(R=256, steps=20, warmup=4, seed=4357787, fp64_value=None):
    """The headline ensemble mode: the replicas in split-precision storage,
    advanced by one batch (HIP-graph replay, autotuned tile -- which sets the
    cells per workgroup of the cell pass -- binning table slots and re-sort
    period).  Timed over at least `steps` intervals, a whole number of re-sort
    periods (one re-sort in each); the kernel means come from a further run
    without the graph (HIP events)."""
    from main_amd.batch import CollectiveBatch
    from main_amd._lib import device_synchronize
    t0 = time.time()
    engines = ensemble_engines(R, seed, True)
    print(f"# split batched: {R} replicas built in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    batch = CollectiveBatch(engines, tile=4096)
    info = batch.info()
    times = batch.autotune(tiles=(2048, 4096, 8192), slots=(1024, 2048), ncoll=10)
    periods = batch.tune_sort_period((10, 20, 40))
    steps = -(-steps // batch.sort_period) * batch.sort_period
    batch.run(warmup)
    device_synchronize()
    t0 = time.perf_counter()
    batch.run(steps)
    device_synchronize()
    dt = time.perf_counter() - t0
    batch.run(16, use_graph=False)
    kms = batch.kernel_ms()
    batch.close()
    for e in engines:
        e.close()
    n = info["particles"]
    ms = dt * 1000.0 / steps
    value = n * steps / dt
    kernels = kernel_rooflines(kms, ALGO_BYTES_SPLIT, n)
    slots = batch.bulk_slots or (2048 if batch.tile >= 4096 else 1024)
    for k, v in kernels.items():   # the binning pass of the tuned table size
        v["traffic"] = read_traffic(f"sb_bulk<false,512,{slots},4>" if k == "sb_bulk" else k, TRAFFIC_SPLIT_FILE)
    print(f"# split batched: {ms:.3f} ms/interval, {value:.3e} particle-updates/s, kernels {kms}", file=sys.stderr,
          flush=True)
    out = {"replicas": R, "particles": n, "collision_cells": info["cells"], "timed_intervals": steps,
           "ms_per_interval": ms,
           "particle_updates_per_s": value, "replicas_per_hour_2000_collisions": R * 3600.0 / (2000 * ms * 1e-3),
           "tile": batch.tile, "bulk_slots": batch.bulk_slots, "sort_every_collisions": batch.sort_period,
           "sort_period_ms_per_interval": periods,
           "autotune_ms_per_interval": {"x".join(map(str, k)): v for k, v in times.items()},
           "roofline_step": step_roofline(SPLIT_STEP_BYTES, n, ms), "kernel_ms": kms,
           "kernels": kernels, "rng": "device Philox4x32-10", "graph": True,
           "dtype": "f32 solvent (cell + offsets), f64 near set and dimers"}
    if fp64_value:
        out["vs_fp64_headline"] = value / fp64_value
    return out


# kernels of an interval without a re-sort (the two-pass collisions) and of a re-sort
STEP_KERNELS = ("co_near_engine", "co_far", "co_zone_mark", "co_collide_far", "co_rotate_far", "co_bin_rest",
                "co_collide_late", "interval_begin", "bits_count", "bits_emit_scan", "co_far_handoff")
SORT_KERNELS = ("sort_bin", "sort_scatter", "sort_order", "sort_gather_six", "sort_gather_one", "sort_gather_bits",
                "co_bulk", "co_scatter", "co_collide", "co_rotate", "rotate_begin", "scan_partial",
                "scan_final_sums")


def step_traffic(sort_every=10):
    """Counter bytes of one interval: every interval's kernels, plus the
    re-sort interval's own launches once per sort_every intervals; None
    unless the committed summary covers the main ones."""
    per = {}
    if not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        for line in f:
            m = re.match(r"traffic_bytes_per_launch\s+(\S+)\s+(\S+)", line)
            if m:   # by base name (co_near_engine<1> -> co_near_engine)
                per.setdefault(m.group(1).split("<")[0], float(m.group(2)))
    if not all(k in per for k in ("co_near_engine", "co_far", "co_collide_far", "co_rotate_far")):
        return None
    return sum(per.get(k, 0.0) for k in STEP_KERNELS) + sum(per.get(k, 0.0) for k in SORT_KERNELS) / sort_every


def read_traffic(kernel, path=TRAFFIC_FILE):
    """Bytes per launch of `kernel` between L2 and memory from a committed
    counter summary (an upper bound on its HBM bytes): the line of that exact
    name (template arguments without spaces, e.g. sb_bulk<false,512,1024,4>),
    else the first line of that base name."""
    if not os.path.exists(path):
        return None
    base = None
    with open(path) as f:
        for line in f:
            m = re.match(r"traffic_bytes_per_launch\s+(\S+)\s+(\S+)", line)
            if not m:
                continue
            if m.group(1) == kernel:
                return float(m.group(2))
            if base is None and m.group(1).split("<")[0] == kernel:
                base = float(m.group(2))
    return base


# Counter passes of the traffic figures (one rocprofv3 --pmc run each; at most
# four TCC counters fit one pass): L2 read requests to the memory fabric by
# size, then write requests.  Bytes = 32 R32 + 64 R64 + 128 R128 and
# 64 W64 + 32 (W - W64).  They count every request that leaves L2 for memory,
# including those the Infinity Cache (MALL) then serves: no counter exposed
# here tells those from HBM accesses, so a figure is an upper bound on the
# kernel's HBM bytes (it equals them once the data exceeds the 256 MB MALL,
# as the copy probe's 256 MiB buffers do).
PMC_READ = ("TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum")
PMC_WRITE = ("TCC_EA0_WRREQ_sum", "TCC_EA0_WRREQ_64B_sum")


def summarize_pmc(read_dir, write_dir):
    """Bytes per dispatch of each kernel between L2 and the memory fabric from
    the two counter passes, checked against the copy probe's known bytes."""
    def load(d, counters):
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            raise SystemExit(f"no counter_collection.csv under {d}")
        per = {}
        for fn in files:
            with open(fn) as f:
                for row in csv.DictReader(f):
                    c = row.get("Counter_Name")
                    if c not in counters:
                        continue
                    name = re.sub(r"^void\s+", "", row["Kernel_Name"].replace("(anonymous namespace)::", ""))
                    name = name.split("(")[0].strip()
                    per.setdefault(name, {}).setdefault(c, []).append(float(row["Counter_Value"]))
        return {k: {c: float(np.mean(v)) for c, v in d.items()} for k, d in per.items()}, \
               {k: max(len(v) for v in d.values()) for k, d in per.items()}
    rd, nr = load(read_dir, PMC_READ)
    wr, _ = load(write_dir, PMC_WRITE)

    def rbytes(k):
        c = rd.get(k, {})
        return 32 * c.get(PMC_READ[0], 0.0) + 64 * c.get(PMC_READ[1], 0.0) + 128 * c.get(PMC_READ[2], 0.0)

    def wbytes(k):
        c = wr.get(k, {})
        w, w64 = c.get(PMC_WRITE[0], 0.0), c.get(PMC_WRITE[1], 0.0)
        return 64 * w64 + 32 * (w - w64)
    probe = [k for k in rd if "probe_copy_kernel" in k]
    if probe:
        pk = probe[0]
        print(f"# copy probe, {PROBE_BYTES} B read and written per launch: counters give "
              f"{rbytes(pk):.4g} B read ({rbytes(pk) / PROBE_BYTES:.3f}x), "
              f"{wbytes(pk):.4g} B written ({wbytes(pk) / PROBE_BYTES:.3f}x)")
    print(f"# {'kernel':44s} {'dispatches':>10s} {'read B':>12s} {'written B':>12s} {'bytes/launch':>13s}")
    for k in sorted(rd, key=lambda k: -(rbytes(k) + wbytes(k))):
        short = re.sub(r"^mpcx::", "", k)
        short = re.sub(r"^(ne_solo|ne_batch)::", "", short)
        rb, wb = rbytes(k), wbytes(k)
        print(f"# {short[:44]:44s} {nr[k]:10d} {rb:12.4g} {wb:12.4g} {rb + wb:13.4g}")
        print(f"traffic_bytes_per_launch {short.replace(' ', '')} {rb + wb:.6g}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--size", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--probe-copy", action="store_true", help="also run the copy probe (counter calibration)")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("READ_DIR", "WRITE_DIR"),
                    help="counter CSVs of the two passes (PMC_READ, PMC_WRITE) -> bytes per launch")
    ap.add_argument("--batched", type=int, default=256,
                    help="replicas of the batched ensemble measured beside the headline (0: skip)")
    ap.add_argument("--batched-only", action="store_true", help="measure only the replica ensemble")
    ap.add_argument("--many-dimers", type=int, default=1, help="measure the many-dimer systems beside the headline")
    ap.add_argument("--many-dimers-only", action="store_true")
    ap.add_argument("--split", type=int, default=1, help="measure the split-precision headline beside the FP64 one")
    ap.add_argument("--split-only", action="store_true")
    ap.add_argument("--split-batched-only", action="store_true", help="measure only the split-precision ensemble")
    args = ap.parse_args()
    if args.summarize_pmc:
        summarize_pmc(*args.summarize_pmc)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1", "n_gpus": args.gpus}))
        sys.exit(2)

    import ctypes as C
    from main_amd import systems as S
    from main_amd._lib import check, device_synchronize, lib
    from main_amd.collective import CollectiveEngine
    from main_amd.tour import cell_tour
    if args.probe_copy:   # the counter runs' reference kernel of known bytes
        ms = C.c_double(0)
        check(lib().mpcx_probe_copy(C.c_int64(PROBE_BYTES), C.c_int32(3), C.byref(ms)), "mpcx_probe_copy")
        print(f"# copy probe: {2 * PROBE_BYTES / (ms.value * 1e-3) / 1e9:.0f} GB/s", file=sys.stderr, flush=True)

    if args.batched_only:
        print(json.dumps({"metric": "replicas/hour (2000-collision runs)", **batched(args.batched, args.steps,
                                                                                   args.warmup)}), flush=True)
        return
    if args.split_batched_only:
        print(json.dumps({"metric": "particle-updates/s (MD+MPC), split-precision batched",
                          **split_batched(args.batched, args.steps, args.warmup)}), flush=True)
        return
    if args.split_only:
        print(json.dumps({"metric": "particle-updates/s (MD+MPC), split precision",
                          **split_headline(args.size, args.steps, args.warmup)}), flush=True)
        return
    if args.many_dimers_only:
        out = {name: many_dimers(name, args.steps, args.warmup) for name in MANY_DIMER}
        out["selkov_50cube"] = selkov(args.steps, args.warmup)
        print(json.dumps({"metric": "particle-updates/s (MD+MPC), many dimers", **out}), flush=True)
        return
    p = headline_params(args.size)
    t0 = time.time()
    s = S.collective_initial_state(p, seed=4357787)
    tour = cell_tour(int(p["box"][0]), int(p["box"][1]), int(p["box"][2]))
    print(f"# init {p['NT']} particles in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    eng = CollectiveEngine(p, rng_mode=1, seed=4357787)
    eng.set_state(s, tour=tour)
    eng.setup(do_sort=True)
    eng.run(args.warmup)
    device_synchronize()
    print(f"# warmup {args.warmup} intervals done", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    eng.run(args.steps)
    device_synchronize()
    dt = time.perf_counter() - t0
    kms = eng.kernel_ms()
    prof = eng.near_profile()
    print("# kernel ms per launch:", kms, file=sys.stderr, flush=True)
    print("# near engine per launch:", prof, file=sys.stderr, flush=True)
    st = eng.get_state()
    eng.close()
    NT = p["NT"]
    value = NT * args.steps / dt
    ms_step = dt * 1000.0 / args.steps
    # roofline of the whole interval (the unit one "launch" of the step is): 100 B per
    # particle-update over the interval's wall time; the HBM-bound kernels beside it
    step = step_roofline(FP64_STEP_BYTES, NT, ms_step)
    kernels = kernel_rooflines(kms, ALGO_BYTES, NT)
    for k, v in kernels.items():
        v["traffic"] = read_traffic(k)
    step.update(kernel="collision interval (all its kernels, both streams)", traffic=step_traffic(),
                traffic_note="bytes between L2 and the memory fabric per interval (read and write requests by "
                             "size; the copy probe reads back at 1.000x), the re-sort's launches counted once per "
                             "sort_step: Infinity-Cache hits included, so an upper bound on HBM bytes",
                kernels=kernels, kernel_ms=kms)
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
        "config": {"workload": f"collective_code plain FP64, {args.size}^3 cells with z walls, 10/cell "
                               f"({NT} solvent), 1 dimer, skin 0.3, tau 0.5 = 50 x 0.01, sort every 10 collisions",
                   "rng": "device Philox4x32-10", "md_steps_per_collision": p["MPCtstep"],
                   "particle_md_updates_per_s": value * p["MPCtstep"], "cstep": st["cstep"]},
        "roofline": step,
    }
    if args.batched > 0:
        out["batched"] = batched(args.batched, steps=10, warmup=2)
        out["split_batched"] = split_batched(args.batched, steps=20, warmup=4, fp64_value=value)
        out["split_batched"]["vs_fp64_batched"] = (out["split_batched"]["particle_updates_per_s"] /
                                                   out["batched"]["particle_updates_per_s"])
    if args.split:
        out["split_precision"] = split_headline(args.size)
    if args.many_dimers:
        out["many_dimers"] = {name: many_dimers(name) for name in MANY_DIMER}
        out["selkov_50cube"] = selkov()
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(args.size)
        except Exception as ex:  # neither CPU build is present
            out["cpu_baseline"] = {"value": None, "error": str(ex)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
