"""Parity cases shared by the GPU tests and __graft_entry__.smoke().

Each case builds a seeded system, runs it through the oracle (the C
restatement of the reference, pinned bit-exact to the reference's own Fortran
modules) and through the HIP engine with the same Mersenne-Twister draws, and
compares the results:
  * integer state (flags, shifted-cell index of every particle, census
    counts, collision-cell counts, inside list, neighbour-list counts and
    members): bit-exact;
  * pair forces and potentials of the list members: within 1e-12 of the
    largest member's;
  * positions and velocities: |gpu - ref| <= 1e-12 * max(|ref|, scale)
    (scale 1 for positions, the thermal speed for velocities) -- the only
    deliberate reordering is the tree sum of the dimer forces.
"""
from __future__ import annotations

import copy

import numpy as np

from main_amd import systems as S
from main_amd.collective import CollectiveEngine, host_draws
from main_amd._lib import MtmodStream
from oracle import oracle as O

REL_TOL = 1e-12


def small_collective_params(**kw):
    base = dict(Lx=16.0, Ly=16.0, Lz=20.0, numdims=1, rhoa=5.0, rhob=0.0, Rc=1.5, Rn=2.0, Dcn=4.0,
                MPCt=0.1, MDt=0.01, skinrad=0.3, sort_step=1000000, k1=0.05, T=1.0 / 6.0)
    base.update(kw)
    return S.collective_params(**base)


def identity_tour(p):
    return np.arange(1, p["vol"] + 1, dtype=np.int64)


def build_collective(p, seed=3, flags_mix=False, mt_seed=4357787):
    s = S.collective_initial_state(p, seed=seed)
    if flags_mix:
        s["flag"][:] = np.random.default_rng(seed + 1).integers(0, 3, p["NT"]).astype(np.int32)
    s["mt"], s["mti"] = O.mtmod_seed(mt_seed)
    return s


def close(a, b, scale):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    tol = REL_TOL * np.maximum(np.abs(b), scale)
    bad = np.abs(a - b) > tol
    err = float(np.max(np.abs(a - b) / np.maximum(np.abs(b), scale))) if a.size else 0.0
    return int(bad.sum()), err


def collective_parity_case(ncoll=1, seed=3, p=None, flags_mix=False, sort_at_setup=True, tour=None,
                           react_box=None, near_mode=None, overlap=True):
    """Run ncoll collision intervals on the oracle and on the GPU engine."""
    p = p or small_collective_params()
    s = build_collective(p, seed=seed, flags_mix=flags_mix)
    if react_box is not None:
        s["react_box"] = np.asarray(react_box, np.int32)
    tour = identity_tour(p) if tour is None else tour
    ref = copy.deepcopy(s)
    O.co_run(ref, 0, do_setup=2 if sort_at_setup else 1, tour=tour)
    O.co_run(ref, ncoll * p["MPCtstep"], tour=tour)

    stream = MtmodStream(mt=s["mt"], mti=s["mti"])
    draws = host_draws(stream, p, ncoll)
    eng = CollectiveEngine(p, rng_mode=0, near_mode=near_mode, overlap=overlap)
    try:
        eng.set_state(s, tour=tour)
        eng.setup(do_sort=sort_at_setup)
        eng.run(ncoll, draws)
        g = eng.get_state()
        cells = eng.get_cells()
        lists = eng.get_lists()
    finally:
        eng.close()

    res = {}
    vscale = np.sqrt(p["T"] / p["Ma"])
    for k in ("x", "y", "z"):
        res[k] = close(g[k], ref[k], 1.0)
    for k in ("vx", "vy", "vz"):
        res[k] = close(g[k], ref[k], vscale)
    for k in ("xd", "yd", "zd"):
        res[k] = close(g[k], ref[k], 1.0)
    for k in ("vxd", "vyd", "vzd"):
        res[k] = close(g[k], ref[k], 1e-3)
    res["flag"] = (int(np.sum(g["flag"] != ref["flag"])), 0.0)
    res["ncn"] = (int(np.sum(cells["ncn"] != ref["ncn"])), 0.0)
    res["tcs"] = (int(np.sum(cells["tcs"] != ref["tcs"])), 0.0)
    for k in ("vxcm", "vycm", "vzcm"):
        res[k] = close(cells[k], ref[k], vscale)
    for k in ("uvx", "uvy", "uvz"):
        res[k] = close(cells[k], ref[k], 1.0)
    for k in ("tA", "tB", "tS"):
        res[k] = (int(np.sum(cells[k][: p["vol"]] != ref[k])), 0.0)
    inside_ref = np.sort(ref["il"][: ref["in_num"]])
    res["inside"] = (0 if np.array_equal(lists["in_list"], inside_ref) else 1, 0.0)
    res["nbnum"] = (int(np.sum(lists["nbnum"] != ref["nbnum"])), 0.0)
    # the neighbour lists themselves (ascending particle numbers) and the pair
    # forces and potentials of their members at the last force evaluation:
    # the valid prefix of each list (the reference leaves earlier builds'
    # entries beyond it)
    nb_bad, fp_res = 0, {k: (0, 0.0) for k in ("fpx", "fpy", "fpz", "pot")}
    for k in range(p["numdims"]):
        for i in range(2):
            n = int(ref["nbnum"][k, i])
            if n != int(lists["nbnum"][k, i]) or not np.array_equal(lists["nb"][k, i, :n], ref["nb"][k, i, :n]):
                nb_bad += 1
                continue
            for f in fp_res:
                scale = float(np.max(np.abs(ref[f][k, i, :n]))) if n else 1.0
                b, e = close(lists[f][k, i, :n], ref[f][k, i, :n], scale)
                fp_res[f] = (fp_res[f][0] + b, max(fp_res[f][1], e))
    res["nb"] = (nb_bad, 0.0)
    res.update(fp_res)
    sc = [(g["cstep"], ref["cstep"]), (g["time_count"], ref["time_count"]),
          (g["time_penetrate"], ref["time_penetrate"])]
    res["scalars"] = (int(sum(a != b for a, b in sc)), 0.0)
    bad = {k: v for k, v in res.items() if v[0] != 0}
    worst = max(v[1] for v in res.values())
    summary = f"NT={p['NT']} ncoll={ncoll} worst_rel={worst:.2e} mismatches={bad if bad else 'none'}"
    return dict(ok=not bad, res=res, summary=summary, gpu=g, ref=ref, cells=cells, lists=lists)
