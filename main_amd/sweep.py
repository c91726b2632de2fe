"""Surrogate-guided parameter sweeps of the single-dimer collective motor.

MPC runs are made in batches on the GPU (`CollectiveBatch`): each replica is
one system of the sweep with its own radii, bond length, interaction energies
and bulk back-reaction rate, advanced past a warm-up and then sampled after
every collision for its time-averaged body-frame B density (axisymmetric
(r, z) grid around the dimer axis, origin at the catalytic sphere, z from the
non-catalytic towards the catalytic sphere) and its propulsion velocity (the
dimer's centre-of-mass velocity along that axis, Output_SelkovDimer.f90:144-184
in the Selkov code is the same quantity).

The surrogate (`Surrogate`, BF16 MFMA forward/backward, Adam on FP32 masters)
maps the 8 inputs to the flattened field plus the velocity.  It is used to
rank candidate parameter points; the points it ranks highest get full MPC
runs, and those runs judge its error: a point whose measured velocity differs
from the prediction by more than the tolerance is reported as a fallback (its
MPC value stands, the surrogate's does not).

Inputs (columns of the parameter arrays), one per replica:
  Rc, Rn   catalytic / non-catalytic sphere radii (the LJ sigmas)
  Dcn      rigid bond length
  EaN, EbN A and B energies on the non-catalytic sphere
  EaC, EbC A and B energies on the catalytic sphere
  k1       bulk B->A rate per cell (reaction_mod.f95:8-80)
The reference converts A to B with probability 1 inside the catalytic shell
(forces_mod.f95:79-81, RPab is read but never used), so the reaction
probability is not an input; the two catalytic-sphere energies are.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import systems as S
from .batch import CollectiveBatch
from .collective import CollectiveEngine
from .surrogate import Surrogate
from .tour import cell_tour

PARAMS = ("Rc", "Rn", "Dcn", "EaN", "EbN", "EaC", "EbC", "k1")

# uniform ranges; Dcn is Rc + Rn + a gap drawn from its range; k1 log-uniform
RANGES = dict(Rc=(1.5, 2.5), Rn=(3.0, 4.2), gap=(0.3, 1.5), EaN=(0.5, 1.5), EbN=(0.1, 1.0),
              EaC=(0.5, 1.5), EbC=(0.5, 1.5), k1=(1e-4, 1e-2))


def sample_params(n: int, rng: np.random.Generator, ranges: dict = RANGES) -> np.ndarray:
    u = lambda k: rng.uniform(*ranges[k], size=n)
    Rc, Rn, gap = u("Rc"), u("Rn"), u("gap")
    lo, hi = ranges["k1"]
    k1 = np.exp(rng.uniform(math.log(lo), math.log(hi), size=n))
    return np.stack([Rc, Rn, Rc + Rn + gap, u("EaN"), u("EbN"), u("EaC"), u("EbC"), k1], axis=1)


class Grid:
    """Body-frame sampling grid: nz bins over [zmin, zmax) along the axis, nr
    bins over [0, rmax) across it."""

    def __init__(self, nr=32, nz=64, rmax=12.0, zmin=-20.0, zmax=12.0):
        self.nr, self.nz, self.rmax, self.zmin, self.zmax = nr, nz, rmax, zmin, zmax

    def kw(self):
        return dict(nr=self.nr, nz=self.nz, rmax=self.rmax, zmin=self.zmin, zmax=self.zmax)

    def bin_volume(self) -> np.ndarray:
        dr, dz = self.rmax / self.nr, (self.zmax - self.zmin) / self.nz
        r = np.arange(self.nr + 1) * dr
        ring = math.pi * (r[1:] ** 2 - r[:-1] ** 2) * dz
        return np.broadcast_to(ring, (self.nz, self.nr))


def system_params(row, L: int, rho: float = 10.0, sort_step: int = 500):
    Rc, Rn, Dcn, EaN, EbN, EaC, EbC, k1 = (float(v) for v in row)
    return S.collective_params(Lx=float(L), Ly=float(L), Lz=float(L), numdims=1, rhoa=rho, Rc=Rc, Rn=Rn, Dcn=Dcn,
                               EaC=EaC, EbC=EbC, EaN=EaN, EbN=EbN, T=1.0 / 6.0, skinrad=0.3, MDt=0.01, MPCt=0.5,
                               sort_step=sort_step, k1=k1)


def mpc_runs(X: np.ndarray, L: int = 24, warm: int = 400, avg: int = 2000, seeds=None, grid: Grid | None = None,
             batch_size: int = 256, rho: float = 10.0, chunk: int = 3000, log=None):
    """Full MPC runs of the parameter rows X in batches: `warm` collisions,
    then `avg` collisions sampled after each.  Returns the B number density
    in units of rho, (n, nz, nr), and the mean propulsion velocity, (n,)."""
    grid = grid or Grid()
    n = X.shape[0]
    seeds = np.arange(n) + 7001 if seeds is None else np.asarray(seeds)
    tour = cell_tour(L, L, L)
    fields = np.zeros((n, grid.nz, grid.nr))
    vel = np.zeros(n)
    vol = grid.bin_volume()
    for b0 in range(0, n, batch_size):
        rows = range(b0, min(n, b0 + batch_size))
        t0 = time.time()
        engines = []
        try:
            for i in rows:
                p = system_params(X[i], L, rho)
                rng_seed = int(seeds[i])
                s = S.collective_initial_state(p, seed=rng_seed)
                e = CollectiveEngine(p, rng_mode=1, seed=rng_seed)
                engines.append(e)
                e.set_state(s, tour=tour)
                e.setup(do_sort=True)
            t1 = time.time()
            batch = CollectiveBatch(engines, tile=2048)
            try:
                batch.run(warm)
                c, v, done = 0.0, 0.0, 0
                while done < avg:
                    m = min(chunk, avg - done)
                    cm, vm = batch.body_average(m, **grid.kw())
                    c, v, done = c + cm * m, v + vm * m, done + m
                    if log and done < avg:
                        log(f"  {done}/{avg} averaged collisions")
                c, v = c / max(avg, 1), v / max(avg, 1)
            finally:
                batch.close()
        finally:
            for e in engines:
                e.close()
        fields[b0:b0 + len(rows)] = c / vol / rho
        vel[b0:b0 + len(rows)] = v
        if log:
            log(f"mpc batch {b0 // batch_size}: {len(rows)} runs of {L}^3 cells, built {t1 - t0:.1f}s, "
                f"{warm}+{avg} collisions {time.time() - t1:.1f}s")
    return fields, vel


class _Member:
    """One network of the ensemble, trained with early stopping."""

    def __init__(self, nf, n_hidden, n_layers, seed):
        self.nf = nf
        self.net = Surrogate(n_in=len(PARAMS), n_hidden=n_hidden, n_layers=n_layers, n_out=nf + 1, seed=seed)
        self.best_epoch = 0
        self.val_curve = []

    def fit(self, F, Y, w, stages, batch, seed, val_frac, check):
        n = len(Y)
        perm = np.random.default_rng(seed).permutation(n)
        nv = int(round(val_frac * n)) if val_frac > 0 else 0
        iv, it = perm[:nv], perm[nv:]
        losses = list(self.net.fit(F[it], Y[it], out_weight=w, batch=batch, epochs=0, lr=0.0))
        Xn = (F - self.net.x_mean) / self.net.x_std
        Yn = (Y - self.net.y_mean) / self.net.y_scale
        wn = w / w.sum()
        best, best_p, ep, k = np.inf, self.net.params(), 0, 0
        for epochs, lr in stages:
            done = 0
            while done < epochs:
                m = min(check, epochs - done)
                losses += list(self.net.train_raw(Xn[it], Yn[it], wn, batch=batch, epochs=m, lr=lr,
                                                  seed=seed * 7919 + k))
                k += 1
                done += m
                ep += m
                if nv:
                    r = self.net.forward_raw(Xn[iv]).astype(np.float64) - Yn[iv]
                    vl = float(np.mean(np.sum(wn * r * r, axis=1)))
                    self.val_curve.append((ep, vl))
                    if vl < best:
                        best, best_p, self.best_epoch = vl, self.net.params(), ep
        if nv:
            self.net.set_params(best_p)
        else:
            self.best_epoch = ep
        return np.asarray(losses)


class SweepSurrogate:
    """The MLP surrogate over the sweep's inputs: an ensemble of networks,
    each predicting the flattened body-frame field followed by the velocity.
    Every member is trained on the GPU with its own initialisation and its own
    held-back validation runs (early stopping); the prediction is the members'
    mean and their spread is the surrogate's own uncertainty."""

    def __init__(self, grid: Grid, n_hidden=256, n_layers=3, v_weight: float | None = None, members: int = 8,
                 seed=1):
        self.grid = grid
        self.nf = grid.nz * grid.nr
        self.members = [_Member(self.nf, n_hidden, n_layers, seed + 101 * m) for m in range(members)]
        # the velocity output weighs as much as a quarter of the field bins unless told otherwise
        self.v_weight = self.nf / 4.0 if v_weight is None else v_weight

    def close(self):
        for m in self.members:
            m.net.close()

    @staticmethod
    def features(X):
        X = np.asarray(X, np.float64).copy()
        X[:, 7] = np.log(X[:, 7])         # the rate enters on a log scale
        return X

    def fit(self, X, fields, vel, stages=((600, 1e-3), (400, 3e-4), (400, 1e-4)), batch=64, seed=0,
            val_frac: float = 0.125, check: int = 20):
        """Train every member; returns the first member's training loss per
        epoch.  Each member holds back `val_frac` of the runs (a different
        draw per member), evaluates its weighted validation loss every
        `check` epochs and keeps the parameters with the lowest one."""
        Y = np.concatenate([fields.reshape(len(vel), -1), vel[:, None]], axis=1)
        w = np.ones(self.nf + 1)
        w[-1] = self.v_weight
        F = self.features(X)
        out = None
        for i, m in enumerate(self.members):
            losses = m.fit(F, Y, w, stages, batch, seed + i, val_frac, check)
            out = losses if out is None else out
        return out

    def predict_members(self, X):
        F = self.features(X)
        return np.stack([m.net.predict(F) for m in self.members])

    def predict(self, X, return_spread: bool = False):
        Ys = self.predict_members(X)
        Y = Ys.mean(axis=0)
        f, v = Y[:, :self.nf].reshape(-1, self.grid.nz, self.grid.nr), Y[:, -1]
        if return_spread:
            return f, v, Ys[:, :, -1].std(axis=0)
        return f, v


def field_error(pred, ref):
    """Relative L2 error of fields, per run."""
    a = pred.reshape(len(pred), -1)
    b = ref.reshape(len(ref), -1)
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


def velocity_error(pred, ref):
    """Velocity error over a set of runs: RMS error divided by the RMS velocity."""
    return float(np.sqrt(np.mean((pred - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))


def select(sur: SweepSurrogate, n_candidates: int, k: int, rng: np.random.Generator, ranges: dict = RANGES):
    """The k points with the largest predicted propulsion velocity among
    n_candidates drawn from the sweep's ranges: (points, predicted velocity,
    ensemble spread of the prediction)."""
    C = sample_params(n_candidates, rng, ranges)
    _, v, sd = sur.predict(C, return_spread=True)
    order = np.argsort(-v)[:k]
    return C[order], v[order], sd[order]


def fallback(pred_v, mpc_v, tol: float):
    """Indices of points whose MPC velocity differs from the prediction by more
    than tol (absolute): for these the MPC result is kept and the surrogate's
    is rejected."""
    return np.nonzero(np.abs(pred_v - mpc_v) > tol)[0]


def quadratic_baseline(Xtr, Vtr, Xte, lam: float = 1e-3):
    """Least-squares fit of the velocity by a full quadratic in the
    standardized inputs (ridge lam), a classical response surface to compare
    the surrogate with."""
    F, Fte = SweepSurrogate.features(Xtr), SweepSurrogate.features(Xte)
    mu, sd = F.mean(axis=0), F.std(axis=0) + 1e-12

    def basis(Z):
        cols = [np.ones(len(Z))] + [Z[:, i] for i in range(Z.shape[1])]
        cols += [Z[:, i] * Z[:, j] for i in range(Z.shape[1]) for j in range(i, Z.shape[1])]
        return np.stack(cols, axis=1)
    A, Ate = basis((F - mu) / sd), basis((Fte - mu) / sd)
    w = np.linalg.solve(A.T @ A + lam * np.eye(A.shape[1]), A.T @ Vtr)
    return Ate @ w


def _noise_corrected(mse, noise_var):
    return math.sqrt(max(mse - noise_var, 0.0))


def study(n_train=512, n_hold=64, repeats=4, n_select=64, n_candidates=8192, L=24, warm=600, avg=12000,
          batch_size=320, seed=2024, grid: Grid | None = None, members=8,
          stages=((600, 1e-3), (400, 3e-4), (400, 1e-4)), log=print):
    """The whole workflow: MPC training runs; held-out points each run
    `repeats` times with different seeds (their spread is the MPC noise, and
    their mean the reference the surrogate is judged against); ensemble
    training on the GPU; held-out errors of the surrogate, of one member
    alone, of the training mean (the method without a surrogate) and of a
    quadratic response surface, each measured against the repeat means and
    corrected for their remaining noise; then surrogate-guided selection of
    n_select points out of n_candidates, checked by full MPC runs against the
    randomly drawn held-out points."""
    grid = grid or Grid()
    rng = np.random.default_rng(seed)
    Xtr = sample_params(n_train, rng)
    Xho = sample_params(n_hold, rng)
    X = np.concatenate([Xtr] + [Xho] * repeats)
    seeds = seed * 1000 + np.arange(len(X))
    t0 = time.time()
    F, V = mpc_runs(X, L=L, warm=warm, avg=avg, seeds=seeds, grid=grid, batch_size=batch_size, log=log)
    t_mpc = time.time() - t0
    Ftr, Vtr = F[:n_train], V[:n_train]
    Fr = F[n_train:].reshape(repeats, n_hold, grid.nz, grid.nr)
    Vr = V[n_train:].reshape(repeats, n_hold)
    Fh, Vh = Fr.mean(axis=0), Vr.mean(axis=0)
    v_var = float(np.mean(Vr.var(axis=0, ddof=1)))                 # single-run velocity noise
    f_var = Fr.reshape(repeats, n_hold, -1).var(axis=0, ddof=1).sum(axis=1)   # per point, summed over bins
    v_rms = float(np.sqrt(np.mean(Vh ** 2)))
    f_norm2 = np.sum(Fh.reshape(n_hold, -1) ** 2, axis=1)

    def vel_err(p):
        mse = float(np.mean((p - Vh) ** 2))
        return dict(measured=math.sqrt(mse) / v_rms, noise_corrected=_noise_corrected(mse, v_var / repeats) / v_rms)

    def fld_err(p):
        d2 = np.sum((p.reshape(n_hold, -1) - Fh.reshape(n_hold, -1)) ** 2, axis=1)
        return dict(measured=float(np.mean(np.sqrt(d2 / f_norm2))),
                    noise_corrected=float(np.mean(np.sqrt(np.maximum(d2 - f_var / repeats, 0) / f_norm2))))

    sur = SweepSurrogate(grid, members=members)
    try:
        t0 = time.time()
        losses = sur.fit(Xtr, Ftr, Vtr, stages=stages)
        t_fit = time.time() - t0
        gemm_ms, gemm_flops = sur.members[0].net.train_stats()
        Fp, Vp, Vsd = sur.predict(Xho, return_spread=True)
        Y0 = sur.predict_members(Xho)[0]
        Vq = quadratic_baseline(Xtr, Vtr, Xho)
        res = dict(
            runs=dict(train=n_train, held_out=n_hold, held_out_repeats=repeats, box=L, warm_collisions=warm,
                      averaged_collisions=avg, mpc_seconds=t_mpc, particles_per_run=int(10 * L ** 3)),
            grid=grid.kw(), inputs=list(PARAMS),
            training=dict(members=members, epochs_per_member=int(sum(e for e, _ in stages)),
                          best_epochs=[m.best_epoch for m in sur.members], first_loss=float(losses[0]),
                          final_loss=float(losses[-1]), seconds=t_fit, member0_last_chunk_gemm_ms=gemm_ms,
                          member0_last_chunk_gemm_flops=gemm_flops),
            mpc_noise=dict(velocity_rms=v_rms, single_run_velocity_sd=math.sqrt(v_var),
                           single_run_velocity_sd_rel=math.sqrt(v_var) / v_rms,
                           repeat_mean_velocity_sd_rel=math.sqrt(v_var / repeats) / v_rms,
                           single_run_field_rel=float(np.mean(np.sqrt(f_var / f_norm2)))),
            held_out_velocity_rel_rms=dict(
                surrogate=vel_err(Vp), one_member=vel_err(Y0[:, -1]), training_mean=vel_err(np.full(n_hold, Vtr.mean())),
                quadratic_response_surface=vel_err(Vq)),
            held_out_field_rel_l2=dict(
                surrogate=fld_err(Fp), one_member=fld_err(Y0[:, :-1]),
                training_mean=fld_err(np.broadcast_to(Ftr.mean(axis=0), Fh.shape))),
            held_out_ensemble_spread_rel=float(np.sqrt(np.mean(Vsd ** 2)) / v_rms))
        # surrogate-guided selection, judged by full MPC runs
        Xs, Vs_pred, Vs_sd = select(sur, n_candidates, n_select, rng)
        Fs, Vs = mpc_runs(Xs, L=L, warm=warm, avg=avg, seeds=seed * 1000 + 50000 + np.arange(n_select), grid=grid,
                          batch_size=batch_size, log=log)
        sd1 = math.sqrt(v_var)
        tol = 2.0 * sd1
        fb = fallback(Vs_pred, Vs, tol)
        mse_s = float(np.mean((Vs_pred - Vs) ** 2))
        res["selection"] = dict(
            candidates=n_candidates, selected=n_select, objective="largest propulsion velocity",
            mpc_mean_velocity_selected=float(Vs.mean()), mpc_mean_velocity_random=float(Vh.mean()),
            mpc_best_velocity_selected=float(Vs.max()), mpc_best_velocity_random=float(Vh.max()),
            mpc_best_velocity_training=float(Vtr.max()),
            predicted_mean_velocity_selected=float(Vs_pred.mean()),
            velocity_rel_rms_selected=dict(measured=math.sqrt(mse_s) / float(np.sqrt(np.mean(Vs ** 2))),
                                           noise_corrected=_noise_corrected(mse_s, v_var)
                                           / float(np.sqrt(np.mean(Vs ** 2)))),
            fallback_tolerance=tol, fallback_points=int(len(fb)),
            mpc_runs_spent=n_train + repeats * n_hold + n_select, mpc_runs_to_screen_all=n_candidates)
        data = dict(X_train=Xtr, F_train=Ftr, V_train=Vtr, X_hold=Xho, F_hold=Fr, V_hold=Vr, F_hold_pred=Fp,
                    V_hold_pred=Vp, V_hold_sd=Vsd, V_hold_quad=Vq, X_sel=Xs, V_sel=Vs, V_sel_pred=Vs_pred,
                    V_sel_sd=Vs_sd, losses=losses)
    finally:
        sur.close()
    return res, data


def main(argv=None):
    import argparse
    import json
    import os
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--train", type=int, default=512)
    ap.add_argument("--hold", type=int, default=64)
    ap.add_argument("--select", type=int, default=64)
    ap.add_argument("--candidates", type=int, default=8192)
    ap.add_argument("--box", type=int, default=24)
    ap.add_argument("--warm", type=int, default=600)
    ap.add_argument("--avg", type=int, default=12000)
    ap.add_argument("--repeats", type=int, default=4)
    ap.add_argument("--members", type=int, default=8)
    ap.add_argument("--batch", type=int, default=320)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--out", default="sweep_out")
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    res, data = study(a.train, a.hold, a.repeats, a.select, a.candidates, a.box, a.warm, a.avg, a.batch, a.seed,
                      members=a.members, log=log)
    os.makedirs(a.out, exist_ok=True)
    with open(os.path.join(a.out, "summary.json"), "w") as f:
        json.dump(res, f, indent=1)
    np.savez_compressed(os.path.join(a.out, "dataset.npz"), **data)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
