"""Parameter derivation and synthetic initial states for the two simulators.

The derived quantities follow the reference drivers exactly:
  collective_code/main_pattern.f95:66-76,117-127 (Vol, Vact, particle counts,
  cell counts, MPCtstep, sphere masses, cutoffs), and
  selkov_code/Main_SelkovDimer.f90:47-105.
Initial states are synthetic (uniform positions outside the sphere cutoffs,
Maxwell-Boltzmann velocities with zero total momentum, all particles A/F),
the same recipe as collective_code/initial_mod.f95:122-345 but drawn from
numpy's generator: the step is what is under test, not the initialiser.
"""
from __future__ import annotations

import math
import numpy as np

PI = 4.0 * math.atan(1.0)
CUTOFF = 2.0 ** (1.0 / 6.0)   # global_mod.f95:138


def collective_params(Lx=30.0, Ly=30.0, Lz=30.0, numdims=1, rhoa=10.0, rhob=0.0, rhos=0.0,
                      Ma=1.0, Mb=1.0, Ms=1.0, Rc=2.0, Rn=4.0, Dcn=6.5,
                      EaC=1.0, EbC=1.0, EsC=1.0, EaN=1.0, EbN=0.1, EsN=1.0,
                      Rw=1.0, Ew=1.0, Rm_delta=1.0, Em=1.0, RPab=1.0, RPba=0.0,
                      T=1.0 / 6.0, skinrad=0.3, a0=1.0, rotang=PI / 2.0, MDt=0.01, MPCt=0.5,
                      sort_step=500, k1=1e-3, react_box_type=7, react_dir=0, max_nb=30000):
    """Parameters of one collective_code system (input.txt order, main_pattern.f95:19-64)."""
    box = np.array([Lx, Ly, Lz], dtype=np.float64)
    vol = int(math.trunc(Lx * Ly * Lz))
    vact = vol - numdims * ((4.0 * PI * Rc ** 3) / 3.0 + (4.0 * PI * Rn ** 3) / 3.0)
    na0 = int(math.floor(rhoa * vact))
    nb0 = int(math.floor(rhob * vact))
    ns0 = int(math.floor(rhob * vact))    # main_pattern.f95:70 uses rhob for S as well
    NT = na0 + nb0 + ns0
    p = dict(box=box, vol=vol, numdims=numdims, NT=NT, Na0=na0, Nb0=nb0, Ns0=ns0,
             Ma=Ma, Mb=Mb, Ms=Ms, Rc=Rc, Rn=Rn, Dcn=Dcn, EaC=EaC, EbC=EbC, EsC=EsC,
             EaN=EaN, EbN=EbN, EsN=EsN, Rw=Rw, Ew=Ew, Rm_delta=Rm_delta, Em=Em, RPab=RPab,
             RPba=RPba, T=T, skinrad=skinrad, a0=a0, rotang=rotang, MDt=MDt, MPCt=MPCt,
             sort_step=sort_step, k1=k1, react_box_type=react_box_type, react_dir=react_dir,
             cutoff=CUTOFF, max_nb=max_nb)
    p["mpcnumcell"] = int(math.trunc((Lx / a0) * (Ly / a0) * (Lz / a0) + (Lx / a0) * (Ly / a0)))
    p["rxnnumcell"] = int(math.trunc((Lx / a0) * (Ly / a0) * (Lz / a0)))
    p["MPCtstep"] = int(math.trunc(MPCt / MDt))
    p["Mc"] = (NT / float(vol)) * ((4.0 * PI) / 3.0) * Rc ** 3
    p["Mn"] = (NT / float(vol)) * ((4.0 * PI) / 3.0) * Rn ** 3
    return p


def _pbc(d, L):
    d = np.where(d > 0.5 * L, d - L, d)
    return np.where(d < -0.5 * L, d + L, d)


def place_dimers(p, rng):
    """Random dimer placement with the wall and overlap rejection of
    initial_mod.f95:80-119 (positions then wrapped in x and y)."""
    box, nd, Dcn, Rc, Rn = p["box"], p["numdims"], p["Dcn"], p["Rc"], p["Rn"]
    cut = p["cutoff"]
    w6 = 3.0 ** float(np.float32(1.0) / np.float32(6.0))
    wc = math.sqrt(((p["Rw"] + Rc) * w6) ** 2)
    wn = math.sqrt(((p["Rw"] + Rn) * w6) ** 2)
    xd = np.zeros((nd, 2), order="F")
    yd = np.zeros((nd, 2), order="F")
    zd = np.zeros((nd, 2), order="F")
    # boxes too low for the reference's wall rejection (it needs
    # (Rw+R)*3^(1/6) + R*2^(1/6) of clearance on both sides of each sphere,
    # and initial_mod.f95:80-119 then never ends): every dimer at mid-height
    # with its bond in the x-y plane, rejecting overlaps only
    low = box[2] < 2.0 * max(wc + Rc * cut, wn + Rn * cut)
    if low and 0.5 * box[2] <= max(Rc, Rn) * cut:
        raise RuntimeError("box too low for a dimer")

    def overlaps(pc, pn, i):
        for j in range(i):
            for (a, ra), (b, rb) in [((pc, Rc), ((xd[j, 0], yd[j, 0], zd[j, 0]), Rc)),
                                     ((pc, Rc), ((xd[j, 1], yd[j, 1], zd[j, 1]), Rn)),
                                     ((pn, Rn), ((xd[j, 0], yd[j, 0], zd[j, 0]), Rc)),
                                     ((pn, Rn), ((xd[j, 1], yd[j, 1], zd[j, 1]), Rn))]:
                d = np.array(a) - np.array(b)
                d[:2] = _pbc(d[:2], box[:2])
                if d @ d < ((ra + rb) * cut + p.get("dimer_gap", 0.0)) ** 2:
                    return True
        return False

    for i in range(nd):
        placed = False
        tries = 0 if low else (20000 if nd > 1 else 2000)
        for _ in range(tries):
            x, y, z, ph, th = rng.random(5)
            ph = 2.0 * PI * ph
            th = 2.0 * th - 1.0
            st = math.sqrt(1.0 - th * th)
            c = np.array([box[0] * x, box[1] * y, box[2] * z])
            u = np.array([st * math.cos(ph), st * math.sin(ph), th])
            pc, pn = c + 0.5 * Dcn * u, c - 0.5 * Dcn * u
            ok = (wc + Rc * cut <= pc[2] <= box[2] - wc - Rc * cut) and \
                 (wn + Rn * cut <= pn[2] <= box[2] - wn - Rn * cut)
            if ok and not overlaps(pc, pn, i):
                placed = True
                break
        if not placed and (low or nd == 1):
            for _ in range(20000):
                x, y, ph = rng.random(3)
                c = np.array([box[0] * x, box[1] * y, 0.5 * box[2]])
                u = np.array([math.cos(2.0 * PI * ph), math.sin(2.0 * PI * ph), 0.0])
                pc, pn = c + 0.5 * Dcn * u, c - 0.5 * Dcn * u
                if not overlaps(pc, pn, i):
                    placed = True
                    break
        if not placed:
            raise RuntimeError("could not place dimer %d" % i)
        pc[:2] -= box[:2] * np.floor(pc[:2] / box[:2])
        pn[:2] -= box[:2] * np.floor(pn[:2] / box[:2])
        xd[i] = [pc[0], pn[0]]
        yd[i] = [pc[1], pn[1]]
        zd[i] = [pc[2], pn[2]]
    return xd, yd, zd


def solvent_outside_spheres(n, box, spheres, rng, periodic_z=False):
    """n uniform positions in the box rejected within any (center, radius^2)."""
    out = np.empty((0, 3))
    while out.shape[0] < n:
        m = int((n - out.shape[0]) * 1.3) + 64
        pos = rng.random((m, 3)) * box
        keep = np.ones(m, bool)
        for (cx, cy, cz), r2 in spheres:
            d = pos - np.array([cx, cy, cz])
            d[:, 0] = _pbc(d[:, 0], box[0])
            d[:, 1] = _pbc(d[:, 1], box[1])
            if periodic_z:
                d[:, 2] = _pbc(d[:, 2], box[2])
            keep &= (d * d).sum(1) >= r2
        out = np.concatenate([out, pos[keep]])
    return out[:n]


def maxwell_velocities(n, kT, m, rng):
    v = rng.standard_normal((n, 3)) * math.sqrt(kT / m)
    return v - v.mean(0)


def collective_dimers(p, seed=4357787, reactive_box=None):
    """Dimer part of the initial state (placement as collective_initial_state)
    for engines that generate the solvent on the device."""
    rng = np.random.default_rng(seed)
    xd, yd, zd = place_dimers(p, rng)
    nd = p["numdims"]
    z = lambda: np.zeros((nd, 2), order="F")
    nrb = int(p["box"][0]) * int(p["box"][1])
    rb = np.zeros(nrb, np.int32) if reactive_box is None else np.asarray(reactive_box, np.int32)
    return dict(xd=xd, yd=yd, zd=zd, vxd=z(), vyd=z(), vzd=z(), fxd=z(), fyd=z(), fzd=z(), react_box=rb)


def collective_initial_state(p, seed=4357787, reactive_box=None):
    """State dict (reference layouts) ready for the oracle or the HIP engine."""
    rng = np.random.default_rng(seed)
    xd, yd, zd = place_dimers(p, rng)
    cC2, cN2 = (p["Rc"] * p["cutoff"]) ** 2, (p["Rn"] * p["cutoff"]) ** 2
    spheres = []
    for k in range(p["numdims"]):
        spheres += [((xd[k, 0], yd[k, 0], zd[k, 0]), cC2), ((xd[k, 1], yd[k, 1], zd[k, 1]), cN2)]
    pos = solvent_outside_spheres(p["NT"], p["box"], spheres, rng)
    vel = maxwell_velocities(p["NT"], p["T"], p["Ma"], rng)
    nd = p["numdims"]
    s = dict(p)
    s.update(x=pos[:, 0].copy(), y=pos[:, 1].copy(), z=pos[:, 2].copy(),
             vx=vel[:, 0].copy(), vy=vel[:, 1].copy(), vz=vel[:, 2].copy(),
             flag=np.zeros(p["NT"], np.int32),
             xd=xd, yd=yd, zd=zd,
             vxd=np.zeros((nd, 2), order="F"), vyd=np.zeros((nd, 2), order="F"),
             vzd=np.zeros((nd, 2), order="F"), fxd=np.zeros((nd, 2), order="F"),
             fyd=np.zeros((nd, 2), order="F"), fzd=np.zeros((nd, 2), order="F"),
             cstep=0, time_count=0, time_penetrate=0, out_num=0, in_num=0)
    nrb = int(p["box"][0]) * int(p["box"][1])
    s["react_box"] = np.zeros(nrb, np.int32) if reactive_box is None else np.asarray(reactive_box, np.int32)
    return s


# ------------------------------------------------------------------ selkov_code
def selkov_params(L=40.0, NF=None, NG=0, NI=None, rho=10.0, Mf=1.0, Mg=1.0, Mi=1.0, Rc=2.0, Rn=4.0,
                  Dcn=7.0, EfC=1.0, EgC=1.0, EiC=1.0, EfN=1.0, EgN=1.0, EiN=1.0, RPfg=1.0, RPgf=0.0,
                  T=1.0, A0=1.0, MDt=0.005, MPCt=0.1, k=(0.0009485, 0.0001, 0.0004, 0.0004, 0.001, 0.0001)):
    """Parameters of one selkov_code system (Input_SelkovDimer.txt order)."""
    box = np.array([L, L, L], np.float64)
    vol = int(L) ** 3
    if NF is None:
        NF = int(0.4 * rho * vol)
    if NI is None:
        NI = int(rho * vol) - NF - NG
    NT = NF + NG + NI
    p = dict(box=box, vol=vol, NT=NT, space=NT, NF=NF, NG=NG, NI=NI, Mf=Mf, Mg=Mg, Mi=Mi, Rc=Rc, Rn=Rn,
             Dcn=Dcn, EfC=EfC, EgC=EgC, EiC=EiC, EfN=EfN, EgN=EgN, EiN=EiN, RPfg=RPfg, RPgf=RPgf,
             T=T / 5.0, A0=A0, MDt=MDt, MPCt=MPCt)   # Main_SelkovDimer.f90:90 divides T by 5
    for i, kv in enumerate(k):
        p["k%d" % (i + 1)] = kv
    p["MPCtstep"] = int(MPCt / MDt)
    p["Mc"] = (NT / float(vol)) * ((4.0 * PI) / 3.0) * Rc ** 3
    p["Mn"] = (NT / float(vol)) * ((4.0 * PI) / 3.0) * Rn ** 3
    p["cutoff"] = CUTOFF
    return p


def selkov_initial_state(p, seed=8154543):
    """Dimer along z at the box centre (Initial_SelkovDimer.f90 places it at the
    centre), solvent outside the cutoffs, F then G then inert flags."""
    rng = np.random.default_rng(seed)
    c = p["box"] / 2.0
    pc = c + np.array([0.0, 0.0, 0.5 * p["Dcn"]])
    pn = c - np.array([0.0, 0.0, 0.5 * p["Dcn"]])
    sph = [(tuple(pc), (p["Rc"] * CUTOFF) ** 2), (tuple(pn), (p["Rn"] * CUTOFF) ** 2)]
    pos = solvent_outside_spheres(p["NT"], p["box"], sph, rng, periodic_z=True)
    vel = maxwell_velocities(p["NT"], p["T"], p["Mf"], rng)
    flag = np.concatenate([np.zeros(p["NF"]), np.ones(p["NG"]), np.full(p["NI"], 3)]).astype(np.int64)
    s = dict(p)
    n = p["NT"]
    s.update(x=pos[:, 0].copy(), y=pos[:, 1].copy(), z=pos[:, 2].copy(),
             vx=vel[:, 0].copy(), vy=vel[:, 1].copy(), vz=vel[:, 2].copy(), flag=flag,
             fx=np.zeros(n), fy=np.zeros(n), fz=np.zeros(n), pfx=np.zeros(n), pfy=np.zeros(n),
             pfz=np.zeros(n))
    dm = np.zeros((2, 12), order="F")
    dm[:, 0], dm[:, 1], dm[:, 2] = [pc[0], pn[0]], [pc[1], pn[1]], [pc[2], pn[2]]
    s["dm"] = dm
    return s
