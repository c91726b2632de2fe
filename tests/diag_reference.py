"""numpy restatement of the diagnostics, used as the checker of the GPU
diagnostics on the same state (test infrastructure).

Every sum runs left to right in particle order (np.cumsum), as the
reference's DO loops add, and every constant has the reference's kind:

* collective observables: collective_code/output_mod.f95:681-756
  (momentum and kinetic energy per sphere, temperatures with the
  default-real literal 2./3.);
* Selkov observables: selkov_code/Output_SelkovDimer.f90:147-300 (the dimer
  enters through its centre-of-mass velocity Vdimav, KEdim is that of the
  centre of mass, species F, G, I are flags 0, 1, 3);
* slabs: output_mod.f95:486-568; shells: Output_Concentration_Catalytic.f90:50-156.

tests/test_diag_pins.py checks the Selkov functions against the reference's
own Output_SelkovDimer/Output_Concentration_Catalytic (oracle/_ref) bit for
bit, the solvent part of the collective function against the same routine
(its solvent loops have the collective's expression shapes), and the
collective dimer terms against exact rational arithmetic.
"""
from __future__ import annotations

import math

import numpy as np

TWO_THIRDS = float(np.float32(2.0) / np.float32(3.0))   # 2./3. in default real = 0.6666666865348816


def seq_sum(a) -> float:
    """0 + a[0] + a[1] + ... left to right (a Fortran accumulation loop)."""
    a = np.asarray(a, dtype=np.float64)
    return float(np.cumsum(np.concatenate([[0.0], a]))[-1]) if a.size else 0.0


def _solvent(st, masses, species):
    """Momenta and kinetic energy of the solvent (output_mod.f95:696-738 /
    Output_SelkovDimer.f90:250-283): p = m v per particle, sums in particle
    order, KE = sum 0.5*(vx px + vy py + vz pz)."""
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
    psol = np.array([seq_sum(px), seq_sum(py), seq_sum(pz)])
    ke = seq_sum(0.5 * ((vx * px + vy * py) + vz * pz))
    counts = [int(np.sum(sp == c)) for c in range(3)]
    pabs = float(np.sum(np.abs(px) + np.abs(py) + np.abs(pz)))
    return psol, ke, counts, pabs


def _bond_axis(xd, nd, box, periodic_z):
    d = np.array(xd[0], dtype=np.float64) - np.array(xd[nd], dtype=np.float64)
    for a in range(3):
        if a < 2 or periodic_z:
            if d[a] > 0.5 * box[a]:
                d[a] -= box[a]
            elif d[a] < -0.5 * box[a]:
                d[a] += box[a]
    ln = math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
    return np.array([d[0] / ln, d[1] / ln, d[2] / ln])


def _v_prop(vd, nd, msph, u):
    """DotVZ of Output_SelkovDimer.f90:168-173 for dimer 0."""
    mc, mn = msph
    vav = [(mc * vd[0][a] + mn * vd[nd][a]) / (mn + mc) for a in range(3)]
    return (vav[0] * u[0] + vav[1] * u[1]) + vav[2] * u[2], vav


def observables(st, masses, species, msph, nd, box, periodic_z, tpot):
    """collective_code/output_mod.f95:681-756 on st (x..vz, flag; xd, vd the
    (2*nd, 3) sphere positions and velocities, sphere s = k + i*nd)."""
    psol, ke, counts, pabs = _solvent(st, masses, species)
    n = sum(counts)
    xd, vd = st["xd"], st["vd"]
    mc, mn = msph
    ptot = psol.copy()
    ked = 0.0
    inv_c, inv_n = 1.0 / mc, 1.0 / mn   # 1./Mc with Mc real(8): a double division
    for k in range(nd):
        pc = [mc * vd[k][a] for a in range(3)]
        pn = [mn * vd[k + nd][a] for a in range(3)]
        for a in range(3):
            ptot[a] = (ptot[a] + pc[a]) + pn[a]
        ked = (ked + (0.5 * inv_c) * ((pc[0] * pc[0] + pc[1] * pc[1]) + pc[2] * pc[2])) + \
            (0.5 * inv_n) * ((pn[0] * pn[0] + pn[1] * pn[1]) + pn[2] * pn[2])
    u = _bond_axis(xd, nd, box, periodic_z)
    vp, _ = _v_prop(vd, nd, msph, u)
    tsol = (ke / float(n)) * TWO_THIRDS if n else 0.0
    return dict(n=n, counts=counts, psol=psol, ptot=ptot, ke_sol=ke, ke_dim=ked, T_sol=tsol,
                T_dim=ked * TWO_THIRDS, E_tot=(ked + ke) + tpot, pabs=pabs, v_prop=float(vp), axis=u)


def selkov_observables(st, masses, msph, box, tpot):
    """selkov_code/Output_SelkovDimer.f90:147-300 on st (flags 0 F, 1 G, 3 I;
    xd, vd the (2, 3) C and N sphere positions and velocities)."""
    psol, ke, counts, pabs = _solvent(st, masses, {0: 0, 1: 1, 3: 2})
    n = sum(counts)
    xd, vd = st["xd"], st["vd"]
    mc, mn = msph
    u = _bond_axis(xd, 1, box, True)
    vp, vav = _v_prop(vd, 1, msph, u)
    pd = [(mc + mn) * vav[a] for a in range(3)]
    ptot = np.array([psol[a] + pd[a] for a in range(3)])
    ked = 0.0 + (0.5 * (1.0 / (mc + mn))) * ((pd[0] * pd[0] + pd[1] * pd[1]) + pd[2] * pd[2])
    ekin = ked + ke
    tsol = (ke / float(n)) * TWO_THIRDS if n else 0.0
    return dict(n=n, counts=counts, psol=psol, ptot=ptot, ke_sol=ke, ke_dim=ked, T_sol=tsol,
                T_dim=ked * TWO_THIRDS, E_tot=ekin + tpot, pabs=pabs, v_prop=float(vp), axis=u)


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


def shell_volumes():
    """ShellVol3..ShellVol20 of Output_Concentration_Catalytic.f90:32-49 (the
    subtracted corrections are default-real literals)."""
    vol = {r: (4.0 / 3.0) * math.pi * float(r) ** 3 for r in (2, 3, 4, 6, 8, 10, 12, 15, 20)}
    f32 = lambda s: float(np.float32(s))   # noqa: E731
    return [vol[3] - vol[2], vol[4] - vol[3] - f32("6.02139"), vol[6] - vol[4] - f32("53.55668"),
            vol[8] - vol[6] - f32("98.4366"), vol[10] - vol[8] - f32("93.051"), vol[12] - vol[10] - f32("17.0169"),
            vol[15] - vol[12], vol[20] - vol[15]]


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
