"""Checks that pin the oracle independently of the reference run:
published known answers and closed forms a plausible slip would break."""
import ctypes as C
import math

import numpy as np

from main_amd import systems as S
from oracle import oracle as O


def test_mt95_known_answer():
    # C++11 [rand.predef]: the 10000th output of std::mt19937 (seed 5489) is 4123659995;
    # mt95's init_by_scalar + genrand_int32 is that generator.
    out = O.mt95_int32(5489, 10000)
    assert int(out[-1]) == 4123659995


def test_mt95_real1_is_shifted_signed_word():
    u = O.mt95_int32(8154543, 50).astype(np.int64)
    a = np.where(u >= 2 ** 31, u - 2 ** 32, u).astype(np.float64)
    expect = a * (1.0 / 4294967295.0) + 2147483648.0 / 4294967295.0
    assert np.array_equal(O.mt95_draws(8154543, 50), expect)


def test_mtmod_twist_is_mt19937():
    # mtmod's twist and tempering equal MT19937: feed its LCG-seeded state to numpy's MT19937
    mt, mti = O.mtmod_seed(4357787)
    bg = np.random.MT19937()
    bg.state = {"bit_generator": "MT19937", "state": {"key": mt.astype(np.uint32), "pos": 624}}
    raw = np.array([bg.random_raw() for _ in range(2000)], dtype=np.float64)
    d, _, _ = O.mtmod_draws(mt, mti, 2000)
    assert np.array_equal(d, raw / 4294967295.0)
    # 1998 seeding: mt(i) = 69069 * mt(i-1) mod 2^32
    assert int(mt[1]) == (69069 * 4357787) % 2 ** 32


def _small():
    p = S.collective_params(Lx=12.0, Ly=12.0, Lz=14.0, rhoa=4.0, Rc=1.5, Rn=2.0, Dcn=4.0, MPCt=0.1)
    s = S.collective_initial_state(p, seed=5)
    return p, s


def test_neighbour_list_brute_force():
    p, s = _small()
    lib = O.oracle_lib()
    nb = np.zeros(p["NT"], np.int64)
    num = C.c_int64(0)
    xd, yd, zd = s["xd"][0, 1], s["yd"][0, 1], s["zd"][0, 1]
    cut = p["Rn"] * p["cutoff"]
    lib.or_co_neighbour_list(C.c_int32(p["NT"]), O._ptr(p["box"]), C.c_double(cut), C.c_double(0.3),
                             O._ptr(s["x"]), O._ptr(s["y"]), O._ptr(s["z"]), C.c_double(xd), C.c_double(yd),
                             C.c_double(zd), C.byref(num), O._ptr(nb), C.c_int64(1))
    d = np.stack([s["x"] - xd, s["y"] - yd, s["z"] - zd], 1)
    for j in range(2):
        L = p["box"][j]
        d[:, j] = np.where(d[:, j] > 0.5 * L, d[:, j] - L, np.where(d[:, j] < -0.5 * L, d[:, j] + L, d[:, j]))
    expect = np.nonzero((d * d).sum(1) <= (cut + 0.3) ** 2)[0] + 1
    assert np.array_equal(nb[: num.value], expect)


def test_wca_force_is_minus_gradient_and_vanishes_at_cutoff():
    lib = O.oracle_lib()
    box = np.array([100.0, 100.0, 100.0])
    R = 2.0
    R6, R12 = R ** 6, R ** 12
    cut2 = (R * 2 ** (1 / 6)) ** 2

    def eval_at(r):
        x = np.array([50.0 + r]); y = np.array([50.0]); z = np.array([50.0])
        flag = np.array([2], np.int32)
        fx, fy, fz, pot = (np.zeros(1) for _ in range(4))
        nb = np.array([1], np.int64)
        lib.or_co_lj_mot_sol(C.c_int32(2), O._ptr(box), O._ptr(flag), O._ptr(x), O._ptr(y), O._ptr(z),
                             C.c_double(50.0), C.c_double(50.0), C.c_double(50.0), O._ptr(fx), O._ptr(fy),
                             O._ptr(fz), O._ptr(pot), C.c_double(0), C.c_double(0), C.c_double(cut2),
                             C.c_double(1.0), C.c_double(1.0), C.c_double(1.0), C.c_double(R6),
                             C.c_double(R12), C.c_double(12 * R12), C.c_double(6 * R6), C.c_int64(1), O._ptr(nb),
                             C.c_int64(1))
        return fx[0], pot[0]

    for r in (2.05, 2.1, 2.2):
        f, _ = eval_at(r)
        e = 1e-6
        dU = (eval_at(r + e)[1] - eval_at(r - e)[1]) / (2 * e)
        assert abs(f + dU) < 1e-6 * max(1.0, abs(f))   # Fpair is the force on the solvent, -dU/dr along +x
    _, u_cut = eval_at(R * 2 ** (1 / 6) * (1 - 1e-12))
    assert abs(u_cut) < 1e-9


def test_mpc_rotation_conserves_cell_momentum_and_energy_and_matches_rodrigues():
    p, s = _small()
    lib = O.oracle_lib()
    NT, mc = p["NT"], p["mpcnumcell"]
    vx0, vy0, vz0 = s["vx"].copy(), s["vy"].copy(), s["vz"].copy()
    vx, vy, vz = vx0.copy(), vy0.copy(), vz0.copy()
    vxcm, vycm, vzcm, tcs, uvx, uvy, uvz = (np.zeros(mc) for _ in range(7))
    ncn = np.zeros(NT, np.int64)
    mt, mti = O.mtmod_seed(99)
    st = O.MtmodState()
    C.memmove(st.mt, mt.ctypes.data, 624 * 4)
    st.mti = mti
    rc = lib.or_co_mpc_rotation(C.c_int32(NT), C.c_double(math.pi / 2), C.c_int64(mc), C.c_double(1.0),
                                O._ptr(p["box"]), O._ptr(s["x"]), O._ptr(s["y"]), O._ptr(s["z"]), O._ptr(vx),
                                O._ptr(vy), O._ptr(vz), O._ptr(vxcm), O._ptr(vycm), O._ptr(vzcm), O._ptr(tcs),
                                O._ptr(ncn), O._ptr(uvx), O._ptr(uvy), O._ptr(uvz), C.byref(st))
    assert rc == 0
    c = ncn - 1
    # cell index by brute force from the same shift
    d, _, _ = O.mtmod_draws(mt, mti, 3)
    sh = 1.0 * (d - 0.5)
    X = s["x"] + sh[0]; X = X - p["box"][0] * np.floor(X / p["box"][0])
    Y = s["y"] + sh[1]; Y = Y - p["box"][1] * np.floor(Y / p["box"][1])
    Z = s["z"] + sh[2]
    fz = np.floor(Z + 1.0) if sh[2] < 0 else np.floor(Z)
    expect = np.floor(X) + p["box"][0] * np.floor(Y) + p["box"][0] * p["box"][1] * fz
    assert np.array_equal(c, expect.astype(np.int64))
    n = np.bincount(c, minlength=mc)
    assert np.array_equal(tcs, n.astype(float))
    for v0, v1, cm in ((vx0, vx, vxcm), (vy0, vy, vycm), (vz0, vz, vzcm)):
        m0 = np.bincount(c, weights=v0, minlength=mc)
        m1 = np.bincount(c, weights=v1, minlength=mc)
        assert np.allclose(m0, m1, rtol=0, atol=1e-12)            # per-cell momentum
        occ = n > 0
        assert np.allclose(cm[occ], m0[occ] / n[occ], rtol=1e-13, atol=1e-15)
    e0 = np.bincount(c, weights=vx0 ** 2 + vy0 ** 2 + vz0 ** 2, minlength=mc)
    e1 = np.bincount(c, weights=vx ** 2 + vy ** 2 + vz ** 2, minlength=mc)
    assert np.allclose(e0, e1, rtol=1e-12, atol=1e-13)           # per-cell kinetic energy
    assert np.allclose(uvx ** 2 + uvy ** 2 + uvz ** 2, 1.0, atol=1e-14)
    # Rodrigues by angle pi/2 about u: R d = d cos + (u x d) sin + u (u.d)(1 - cos)
    i = 7
    u = np.array([uvx[c[i]], uvy[c[i]], uvz[c[i]]])
    cm = np.array([vxcm[c[i]], vycm[c[i]], vzcm[c[i]]])
    dv = np.array([vx0[i], vy0[i], vz0[i]]) - cm
    a = math.pi / 2
    rot = dv * math.cos(a) + np.cross(u, dv) * math.sin(a) + u * (u @ dv) * (1 - math.cos(a))
    assert np.allclose(np.array([vx[i], vy[i], vz[i]]) - cm, rot, atol=1e-14)


def test_rattle_holds_bond_after_steps():
    p, s = _small()
    s["mt"], s["mti"] = O.mtmod_seed(4357787)
    O.co_run(s, 0, do_setup=1)
    O.co_run(s, 25)
    d = np.array([s["xd"][0, 0] - s["xd"][0, 1], s["yd"][0, 0] - s["yd"][0, 1], s["zd"][0, 0] - s["zd"][0, 1]])
    for j in range(2):
        L = p["box"][j]
        d[j] = d[j] - L if d[j] > 0.5 * L else d[j] + L if d[j] < -0.5 * L else d[j]
    assert abs(d @ d - p["Dcn"] ** 2) < 1e-6                     # evolve_mod.f95:388
    dv = np.array([s["vxd"][0, 0] - s["vxd"][0, 1], s["vyd"][0, 0] - s["vyd"][0, 1],
                   s["vzd"][0, 0] - s["vzd"][0, 1]])
    assert abs(2.0 * dv @ d) < 1e-8                                # evolve_mod.f95:710


def test_reaction_threshold_semantics():
    lib = O.oracle_lib()
    vol = 4
    tB = np.array([0, 3, 5, 1], np.int64)
    sB = np.zeros((vol, 100), np.int64, order="F")
    sB[:, 0] = [1, 2, 3, 4]
    flag = np.ones(4, np.int32)
    mt, mti = O.mtmod_seed(7)
    u, _, _ = O.mtmod_draws(mt, mti, 4)
    k1, tau = 0.5, 0.5
    prob = np.array([1 - math.exp(-(k1 * n * tau)) if n else 0.0 for n in tB])
    st = O.MtmodState()
    C.memmove(st.mt, mt.ctypes.data, 624 * 4)
    st.mti = mti
    lib.or_co_reaction_simple(C.c_double(tau), C.c_int64(vol), C.c_double(k1), O._ptr(flag), O._ptr(tB),
                              O._ptr(sB), C.byref(st))
    expect = np.where((u > 0) & (u <= prob), 0, 1)
    assert np.array_equal(flag, expect)
