"""numpy restatement of the diagnostics, used as the checker of the GPU
diagnostics on the same state (test infrastructure).

observables: collective_code/output_mod.f95:681-805 and
selkov_code/Output_SelkovDimer.f90:147-173; slabs: output_mod.f95:486-568;
shells: selkov_code/Output_Concentration_Catalytic.f90:50-156.
"""
from __future__ import annotations

import math

import numpy as np


def observables(st, masses, species, msph, nd, box, periodic_z, tpot):
    f = np.asarray(st["flag"]).astype(np.int64)
    sp = np.full(f.shape, -1)
    for val, col in species.items():
        sp[f == val] = col
    m = np.zeros(f.shape)
    for val, mv in masses.items():
        m[f == val] = mv
    sel = sp >= 0
    vx, vy, vz = (np.asarray(st[k])[sel] for k in ("vx", "vy", "vz"))
    mm = m[sel]
    px, py, pz = mm * vx, mm * vy, mm * vz
    ke = float(np.sum(0.5 * (vx * px + vy * py + vz * pz)))
    counts = [int(np.sum(sp == c)) for c in range(3)]
    n = sum(counts)
    out = dict(n=n, counts=counts, psol=np.array([px.sum(), py.sum(), pz.sum()]), ke_sol=ke,
               T_sol=(ke / n) * (2.0 / 3.0), pabs=float(np.sum(np.abs(px) + np.abs(py) + np.abs(pz))))
    xd, vd = st["xd"], st["vd"]   # (2*nd, 3) positions and velocities, s = k + i*nd
    ked = 0.0
    ptot = out["psol"].copy()
    for k in range(nd):
        pc = msph[0] * vd[k]
        pn = msph[1] * vd[k + nd]
        ptot = ptot + pc + pn
        ked += 0.5 / msph[0] * float(pc @ pc) + 0.5 / msph[1] * float(pn @ pn)
    out.update(ptot=ptot, ke_dim=ked, T_dim=ked * (2.0 / 3.0), E_tot=ke + ked + tpot)
    d = xd[0] - xd[nd]
    for a in range(3):
        if a < 2 or periodic_z:
            if d[a] > 0.5 * box[a]:
                d[a] -= box[a]
            elif d[a] < -0.5 * box[a]:
                d[a] += box[a]
    u = d / math.sqrt(float(d @ d))
    vcm = (msph[0] * vd[0] + msph[1] * vd[nd]) / (msph[1] + msph[0])
    out["v_prop"] = float(vcm @ u)
    out["axis"] = u
    return out


def slabs(st, box, xs, ys, zs, half, w):
    f = np.asarray(st["flag"])
    x, y, z = (np.asarray(st[k]) for k in ("x", "y", "z"))
    nx, ny, nz = (int(b / w) for b in box)
    jx, jy, jz = (np.ceil(a / w).astype(np.int64) - 1 for a in (x, y, z))
    out = {}
    for s, name in ((0, "A"), (1, "B")):
        ok = f == s
        zsl = ok & (z <= zs + half) & (z >= zs - half)
        bx = (jx >= 0) & (jx < nx)
        by = (jy >= 0) & (jy < ny)
        bz = (jz >= 0) & (jz < nz)
        out[name + "1d"] = np.bincount(jx[zsl & bx], minlength=nx)[:nx]
        a = np.zeros((nx, ny), np.int64)
        np.add.at(a, (jx[zsl & bx & by], jy[zsl & bx & by]), 1)
        out[name + "xy"] = a
        ysl = ok & (y <= ys + half) & (y >= ys - half) & bx & bz
        a = np.zeros((nx, nz), np.int64)
        np.add.at(a, (jx[ysl], jz[ysl]), 1)
        out[name + "xz"] = a
        xsl = ok & (x <= xs + half) & (x >= xs - half) & by & bz
        a = np.zeros((ny, nz), np.int64)
        np.add.at(a, (jy[xsl], jz[xsl]), 1)
        out[name + "yz"] = a
    return out


def shells(st, box, c):
    f = np.asarray(st["flag"])
    hb = 0.5 * np.asarray(box)
    R = []
    for a, k in enumerate(("x", "y", "z")):
        X = np.asarray(st[k]) + (hb[a] - c[a])
        X = np.where(X > box[a], X - box[a], np.where(X < 0.0, X + box[a], X))
        R.append(X - hb[a])
    R2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2]
    lim = np.array([4.0, 9.0, 16.0, 36.0, 64.0, 100.0, 144.0, 225.0, 400.0])
    out = np.zeros((3, 8), np.int64)
    for col, val in enumerate((0, 1, 3)):
        r = R2[(f == val) & (R2 >= 4.0) & (R2 < 400.0)]
        b = np.searchsorted(lim, r, side="right") - 1
        out[col] = np.bincount(b, minlength=8)[:8]
    return out


def body_density(st, species, box, periodic_z, origin, axis, nr, nz, rmax, zmin, zmax):
    f = np.asarray(st["flag"]).astype(np.int64)
    sp = np.full(f.shape, -1)
    for val, col in species.items():
        sp[f == val] = col
    d = []
    for a, k in enumerate(("x", "y", "z")):
        v = np.asarray(st[k]) - origin[a]
        if a < 2 or periodic_z:
            v = np.where(v > 0.5 * box[a], v - box[a], np.where(v < -0.5 * box[a], v + box[a], v))
        d.append(v)
    zr = (d[0] * axis[0] + d[1] * axis[1]) + d[2] * axis[2]
    p = [d[a] - zr * axis[a] for a in range(3)]
    rr = np.sqrt((p[0] * p[0] + p[1] * p[1]) + p[2] * p[2])
    ok = (sp >= 0) & (zr >= zmin) & (zr < zmax) & (rr < rmax)
    dr, dz = rmax / nr, (zmax - zmin) / nz
    iz = np.minimum(np.floor((zr[ok] - zmin) / dz).astype(np.int64), nz - 1)
    ir = np.minimum(np.floor(rr[ok] / dr).astype(np.int64), nr - 1)
    out = np.zeros((3, nz, nr), np.int64)
    np.add.at(out, (sp[ok], iz, ir), 1)
    return out
