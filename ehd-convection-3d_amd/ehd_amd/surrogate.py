"""Learned warm starts: data generation with the batched solver, FNO training
on the GPU, and the steps-to-steady evaluation against cold starts.

Protocol (the acceptance rule of the surrogate):
  * training states: boxes with (T, C) drawn from the training ranges, run
    from the hydrostatic start plus seeded velocity noise (amplitude 1e-3)
    to a steady state (|dNu| < tol over `check_every` steps) in batches;
    each state is put in a canonical position (`canonicalize`): from noise a
    box picks its own pattern position (and, for rolls, direction), which the
    network cannot know;
  * targets: the steady u, v, w in units of the drift velocity K V / Lz and
    the deviations of q and phi from the hydrostatic profile in units of
    0.2 q0 and 0.05 V (the injected charge and the voltage, scaled so that
    every channel has an rms of about 0.5 and weighs alike in the loss);
  * evaluation: for held-out (T, C) draws, a cold box (hydrostatic start +
    noise) and a warm box run to steady (the cold boxes in one batch, the
    warm ones in another).  The warm box starts
    from the network's fields by the consistent start (`ehd_init_consistent`:
    a few thousand steps that collide towards the predicted velocity and
    charge, building the populations' non-equilibrium parts; an equilibrium
    start from even the exact steady fields needs ~5e4 steps to satisfy the
    detector again).  Those steps count for both boxes; a warm start is accepted if its final Nu is within
    0.5 % of the cold start's, otherwise the cold start is the answer.  The
    gain is the ratio of the steps to steady.

Usage (one MI355X):  python -m ehd_amd.surrogate --train 64 --test 32
prints one JSON summary line.

Grid, time step and range.  The boxes are 64x64x64 (dz = 1/63, lateral
period 1.016 Lz, close to the critical wavelength of the instability), run
with dt/dx = 0.04 (the reference's constant is 0.01: the drift time Lz^2/(K V)
is then 25,200 steps on this grid, and steady states take millions of steps;
at 0.04 it is 6,300 steps).  T is drawn from 200..300 and C from 5..20: there
the flows convect (Nu 1.4-1.9), settle from noise to steady square cells in
~1.7-2.1e5 steps and the lattice stays stable.  On coarser z grids (dz = 1/31) or at higher T
the scheme (the same in FP64 and FP32S, with the charge relaxation
parameter 1e-6 of the reference) goes non-finite within 1e5 steps, and at
T >~ 500 the flows are not steady.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import (FP32S, INIT_FIELDS, INIT_HYDROSTATIC, PERTURB_NOISE, Batch, Box, EhdError, derive, make_params)
from .fno import FNO

D_IN, D_OUT = 6, 5
CFL = 0.04          # dt/dx of the surrogate's boxes
# units of the q and phi targets, in q0 and V: the deviations from the
# hydrostatic profile are ~0.1 q0 and ~0.025 V where the velocities are
# ~0.5-1.4 drift velocities
Q_UNIT, PHI_UNIT = 0.2, 0.05


def box_params(n, T, C, M=10.0, seed=1, init=INIT_HYDROSTATIC, amp=1e-3, cfl=CFL, pattern=PERTURB_NOISE):
    nx, ny, nz = n
    return make_params(nx, ny, nz, T=T, C_=C, M=M, Fe=4000.0, Pr=10.0, lam=0.25, precision=FP32S,
                       init_mode=init, seed=seed, amp=amp, cfl=cfl, pattern=pattern)


def features(n, T, C, M=10.0):
    nx, ny, nz = n
    a = np.empty((D_IN, nz, ny, nx), np.float32)
    a[0] = math.log(T / 400.0)
    a[1] = math.log(C / 10.0)
    a[2] = math.log(M / 10.0)
    a[3] = (np.arange(nx) / nx)[None, None, :]
    a[4] = (np.arange(ny) / ny)[None, :, None]
    a[5] = (np.arange(nz) / (nz - 1))[:, None, None]
    return a


def hydrostatic_profile(p):
    """q(z), phi(z) of the analytic hydrostatic state of the box (what the
    hydrostatic start holds), read back from a freshly initialised box."""
    q = make_params(p.nx, p.ny, p.nz, T=p.T, C_=p.C, M=p.M, Fe=p.Fe, lam=p.lambda_f, precision=p.precision,
                    init_mode=INIT_HYDROSTATIC, seed=1, amp=0.0, cfl=p.cfl)
    with Box(q) as b:
        b.init()
        f = b.read_fields(("q", "phi"))
    return f["q"][:, 0, 0].copy(), f["phi"][:, 0, 0].copy()


def scales(p):
    d = derive(p)
    return d.K * p.voltage / p.lz, d.charge0, p.voltage


def targets_from_fields(p, f, prof):
    U, Q, V = scales(p)
    qz, phz = prof
    t = np.empty((D_OUT,) + f["ux"].shape, np.float32)
    t[0], t[1], t[2] = f["ux"] / U, f["uy"] / U, f["uz"] / U
    t[3] = (f["q"] - qz[:, None, None]) / (Q_UNIT * Q)
    t[4] = (f["phi"] - phz[:, None, None]) / (PHI_UNIT * V)
    # wall rows of the macroscopic outputs are raw collide values: use the extrapolation every consumer uses
    for k in range(4):
        t[k, 0] = 2 * t[k, 1] - t[k, 2]
        t[k, -1] = 2 * t[k, -2] - t[k, -3]
    return t


def _shift(t, axis, d):
    """Periodic shift by d nodes along an axis (t(x - d)), spectral, so d may
    be fractional; exact for the band-limited fields of the lattice."""
    n = t.shape[axis]
    k = np.fft.rfftfreq(n) * n
    ph = np.exp(-2j * np.pi * k * d / n)
    shape = [1] * t.ndim
    shape[axis] = ph.size
    return np.fft.irfft(np.fft.rfft(t, axis=axis) * ph.reshape(shape), n=n, axis=axis)


def _dominant_modes(w):
    """The strongest Fourier mode (ky, kx) of a [Y][X] plane and the strongest
    one not parallel to it (None if it is weaker than 5 % of the first), with
    their phases; one mode of each conjugate pair."""
    Y, X = w.shape
    F = np.fft.fft2(w)
    A = np.abs(F)
    A[0, 0] = 0.0
    order = np.argsort(A.ravel())[::-1]
    first = None
    for i in order[:64]:
        iy, ix = divmod(int(i), X)
        ky = iy if iy <= Y // 2 else iy - Y
        kx = ix if ix <= X // 2 else ix - X
        if first is None:
            first = (ky, kx, float(np.angle(F[iy, ix])), float(A[iy, ix]))
            continue
        if A[iy, ix] < 0.05 * first[3]:
            break
        if ky * first[1] - kx * first[0] != 0:
            return first, (ky, kx, float(np.angle(F[iy, ix])), float(A[iy, ix]))
    return first, None


def canonicalize(t):
    """Remove the translation (and, on square boxes, the x/y exchange)
    symmetry of a steady convection pattern: every periodic shift of a steady
    state is a steady state, and the network cannot know which position a run
    picks.  With (ky1, kx1) the strongest Fourier mode of w at mid-height and
    (ky2, kx2) the strongest one not parallel to it, the field is shifted
    (fractional spectral shift of every channel) so that both have phase
    zero: kx a + ky b = -phase / 2 pi for both, a = x0/X, b = y0/Y.  When
    several shifts do that (|det| > 1, or a single mode with |k| > 1), the one
    that puts the largest upflow w at the origin is taken.  A single mode
    (rolls) fixes the position across the rolls only.  On square boxes x and
    y (and u and v) are first exchanged if the strongest mode points more
    along y than along x.  t: [5, Z, Y, X] (u, v, w, q, phi) -> the
    canonical copy (float64)."""
    t = np.array(t, dtype=np.float64, copy=True)
    Z, Y, X = t.shape[1:]
    m1, m2 = _dominant_modes(t[2, Z // 2])
    if m1 is None or m1[3] == 0.0:
        return t
    if X == Y and abs(m1[0]) > abs(m1[1]):
        t = t.swapaxes(2, 3).copy()
        t[[0, 1]] = t[[1, 0]]
        m1, m2 = _dominant_modes(t[2, Z // 2])
    F = np.fft.fft2(t[2, Z // 2])
    fy = (np.fft.fftfreq(Y) * Y)[:, None]
    fx = (np.fft.fftfreq(X) * X)[None, :]

    def w_at(a, b):   # w at x = a X, y = b Y
        return float(np.real(np.sum(F * np.exp(2j * np.pi * (fy * b + fx * a)))))

    ky1, kx1, p1, _ = m1
    if m2 is None:
        if kx1 != 0:
            cands = [((-p1 / (2 * np.pi) + n) / kx1, 0.0) for n in range(abs(kx1))]
        else:
            cands = [(0.0, (-p1 / (2 * np.pi) + n) / ky1) for n in range(abs(ky1))]
    else:
        ky2, kx2, p2, _ = m2
        M = np.array([[kx1, ky1], [kx2, ky2]], float)
        d = int(round(abs(np.linalg.det(M))))
        base = np.linalg.solve(M, np.array([-p1, -p2]) / (2 * np.pi))
        cands = [tuple(base + np.linalg.solve(M, np.array([n1, n2], float)))
                 for n1 in range(d) for n2 in range(d)]
    a, b = max(cands, key=lambda c: w_at(*c))
    t = _shift(t, 3, -a * X)
    t = _shift(t, 2, -b * Y)
    return t


def fields9_from_prediction(p, y, prof, noise_seed=None, amp=0.0):
    """ehd_init fields (rho, q, phi, u, v, w, Ex, Ey, Ez) from a network output."""
    U, Q, V = scales(p)
    d = derive(p)
    qz, phz = prof
    nz, ny, nx = p.nz, p.ny, p.nx
    f = np.zeros((9, nz, ny, nx))
    f[0] = p.rho0
    f[1] = qz[:, None, None] + Q_UNIT * Q * y[3]
    # The injection wall takes the injected charge, not a prediction: its
    # charge populations are rebuilt from their own previous values every step
    # (LBM.cu:1281-1353), so q(z=0) follows q' = 2 q_inj - q and a start away
    # from q_inj oscillates with period 2 for ever and shifts the steady state.
    f[1][0] = qz[0]
    f[2] = phz[:, None, None] + PHI_UNIT * V * y[4]
    f[2][0] = p.voltage
    f[2][-1] = 0.0
    f[3], f[4], f[5] = U * y[0], U * y[1], U * y[2]
    if noise_seed is not None and amp > 0:
        rng = np.random.default_rng(noise_seed)
        f[3:6] += amp * rng.uniform(-1, 1, (3, nz, ny, nx))
    f[3:6, 0] = 0.0
    f[3:6, -1] = 0.0
    phi = f[2]
    f[6] = (np.roll(phi, 1, axis=2) - np.roll(phi, -1, axis=2)) / (2 * d.dx)
    f[7] = (np.roll(phi, 1, axis=1) - np.roll(phi, -1, axis=1)) / (2 * d.dy)
    ez = np.empty_like(phi)
    ez[1:-1] = (phi[:-2] - phi[2:]) / (2 * d.dz)
    ez[0], ez[-1] = ez[1], ez[-2]
    f[8] = ez
    return f


class Steady:
    def __init__(self, steps, converged, nusselt):
        self.steps, self.converged, self.nusselt = steps, converged, nusselt
        self.failed = False      # the box went non-finite (the lattice was unstable for it)


def _diag(box):
    try:
        return box.diagnostics()
    except EhdError as e:
        if e.code == -5:     # EHD_ENONFINITE
            return None
        raise


def run_batch_to_steady(params, fields=None, check_every=1000, tol=1e-5, max_steps=200000, min_steps=0,
                        confirm=2, tag="", consistent=0):
    """Run boxes to steady state in one batch; returns per box (steps,
    converged, Nu) and the final fields.  A box is steady once, over
    `confirm` consecutive windows of check_every steps, both |dNu| < tol and
    |d umax| < tol * max(umax, 1e-6 K V/Lz): near the hydrostatic start Nu
    hardly moves while the instability grows, but umax changes by its growth
    rate, so the start cannot pass for steady.  min_steps (optional) delays
    the test.  consistent > 0: the boxes with fields start by
    `ehd_batch_init_consistent` with that many iterations (every box then
    has taken that many steps, and they count).  Progress goes to stderr
    every half minute."""
    n = len(params)
    ufloor = [1e-6 * scales(p)[0] for p in params]
    with Batch(params) as bt:
        if consistent > 0 and fields is not None:
            bt.init_consistent(fields, consistent)
            taken = consistent
        else:
            bt.init(fields)
            taken = 0
        d0 = [_diag(b) for b in bt.boxes]
        prev = [d.nusselt if d else math.nan for d in d0]
        prevu = [d.umax if d else math.nan for d in d0]
        res = [Steady(max_steps, False, v) for v in prev]
        for r, d in zip(res, d0):
            if d is None:
                r.failed, r.steps = True, taken
        quiet = [0] * n
        t0, tp = time.perf_counter(), time.perf_counter()
        while taken < max_steps and not all(r.converged or r.failed for r in res):
            k = min(check_every, max_steps - taken)
            bt.step(k)
            taken += k
            for i, b in enumerate(bt.boxes):
                r = res[i]
                if r.converged or r.failed:
                    continue
                d = _diag(b)
                if d is None:
                    r.failed, r.steps = True, taken
                    continue
                r.nusselt = d.nusselt
                small = (k == check_every and abs(d.nusselt - prev[i]) < tol
                         and abs(d.umax - prevu[i]) < tol * max(d.umax, ufloor[i]))
                quiet[i] = quiet[i] + 1 if small else 0
                if quiet[i] >= confirm and taken >= min_steps:
                    r.converged, r.steps = True, taken
                prev[i] = d.nusselt
                prevu[i] = d.umax
            if time.perf_counter() - tp > 30:
                tp = time.perf_counter()
                print(f"{tag} {taken} steps, {sum(r.converged for r in res)}/{n} steady "
                      f"({tp - t0:.0f} s)", file=sys.stderr, flush=True)
        out = [b.read_fields(("q", "phi", "ux", "uy", "uz")) for b in bt.boxes]
    return res, out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, nargs=3, default=(64, 64, 64))
    ap.add_argument("--train", type=int, default=64)
    ap.add_argument("--test", type=int, default=32)
    ap.add_argument("--batch-boxes", type=int, default=64, help="boxes per solver batch")
    ap.add_argument("--check-every", type=int, default=2000)
    ap.add_argument("--tol", type=float, default=1e-5)
    ap.add_argument("--max-steps", type=int, default=300000)
    ap.add_argument("--gen-max-steps", type=int, default=None,
                    help="step limit of the training-state generation (default: --max-steps)")
    ap.add_argument("--T-range", type=float, nargs=2, default=(200.0, 300.0),
                    help="training and test range of T (steady convection on this grid)")
    ap.add_argument("--C-range", type=float, nargs=2, default=(5.0, 20.0))
    ap.add_argument("--min-steps", type=int, default=0)
    ap.add_argument("--epochs", type=int, default=1000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--width", type=int, default=32)
    ap.add_argument("--modes", type=int, nargs=3, default=(8, 8, 16), help="Fourier modes along x, y, z")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--save-net", default="", help="save the trained parameters (.npy)")
    ap.add_argument("--test-seed", type=int, default=None, help="draw the held-out (T, C) from their own generator")
    ap.add_argument("--test-offset", type=int, default=0, help="skip this many draws of that generator")
    ap.add_argument("--load-net", default="", help="skip generation and training: evaluate these parameters")
    ap.add_argument("--consistent", type=int, default=5000,
                    help="consistent-start iterations of the warm boxes (0: equilibrium start); counted as steps "
                         "of every box of the evaluation batch")
    args = ap.parse_args(argv)
    n = tuple(args.grid)
    rng = np.random.default_rng(args.seed)
    t0 = time.perf_counter()

    def draws(k):
        t0_, t1_ = args.T_range
        c0_, c1_ = args.C_range
        return [(float(rng.uniform(t0_, t1_)), float(math.exp(rng.uniform(math.log(c0_), math.log(c1_)))))
                for _ in range(k)]

    train_draws = draws(args.train)
    if args.test_seed is None:
        test_draws = draws(args.test)
    else:
        # held-out draws from their own generator: several runs can evaluate
        # disjoint slices of one draw sequence against the same (deterministic) training
        rng_t = np.random.default_rng(args.test_seed)
        t0_, t1_ = args.T_range
        c0_, c1_ = args.C_range
        seq = [(float(rng_t.uniform(t0_, t1_)), float(math.exp(rng_t.uniform(math.log(c0_), math.log(c1_)))))
               for _ in range(args.test_offset + args.test)]
        test_draws = seq[args.test_offset:]
    profs = {}

    def prof_of(p):
        key = (round(p.C, 12),)
        if key not in profs:
            profs[key] = hydrostatic_profile(p)
        return profs[key]

    net = FNO(n[0], n[1], n[2], width=args.width, modes=tuple(args.modes), max_batch=8)
    X, gen_steps, losses, ntr, t_gen, t_train = None, [], [], 0, 0.0, 0.0
    if args.load_net:
        # a network trained by an earlier run with the same grid, width and modes
        net.set_params(np.load(args.load_net))
    else:
        # 1. training states from the batched solver
        X, Y = [], []
        gen_steps = []
        for s in range(0, len(train_draws), args.batch_boxes):
            chunk = train_draws[s:s + args.batch_boxes]
            params = [box_params(n, T, C, seed=1000 + s + i) for i, (T, C) in enumerate(chunk)]
            res, fields = run_batch_to_steady(params, check_every=args.check_every, tol=args.tol,
                                              max_steps=args.gen_max_steps or args.max_steps,
                                              min_steps=args.min_steps,
                                              tag=f"train batch {s // args.batch_boxes}:")
            for p, (T, C), f, r in zip(params, chunk, fields, res):
                if not r.converged or not (0.5 < r.nusselt < 5.0):
                    continue     # not steady (or the lattice went unstable): not a training state
                X.append(features(n, T, C))
                Y.append(canonicalize(targets_from_fields(p, f, prof_of(p))))
                gen_steps.append(r.steps)
            print(f"generated {len(X)} states ({time.perf_counter() - t0:.0f} s)", file=sys.stderr, flush=True)
        X = np.stack(X)
        Y = np.stack(Y)
        t_gen = time.perf_counter() - t0

        # 2. training on the GPU (the library's forward/backward kernels, Adam)
        net.set_params(net.init_params(seed=args.seed))
        net.set_dataset(X, Y)        # resident in HBM: mini-batches are gathered on the device
        ntr = len(X)
        bs = 8
        steps_per_epoch = max(1, ntr // bs)
        total = args.epochs * steps_per_epoch
        losses = []
        t1 = time.perf_counter()
        k = 0
        for ep in range(args.epochs):
            order = rng.permutation(ntr)
            el = []
            for j in range(steps_per_epoch):
                idx = order[j * bs:(j + 1) * bs]
                lr = args.lr * 0.5 * (1 + math.cos(math.pi * k / total))
                el.append(net.train_step_indices(idx, lr))
                k += 1
            losses.append(float(np.mean(el)))
            if ep % 5 == 0 or ep == args.epochs - 1:
                print(f"epoch {ep}: loss {losses[-1]:.4e} ({time.perf_counter() - t1:.0f} s)", file=sys.stderr, flush=True)
        t_train = time.perf_counter() - t1
        if args.save_net:
            np.save(args.save_net, net.get_params())

    if not test_draws:
        print(json.dumps({"grid": list(n), "train_states": int(ntr), "generation_s": round(t_gen, 1),
                          "training_s": round(t_train, 1), "final_train_loss": losses[-1] if losses else None,
                          "saved": args.save_net or None}), flush=True)
        return 0

    # 3. steps to steady: cold (hydrostatic + noise) and warm (the network's fields), each
    # set in a batch of its own (a batch steps until its slowest box is steady)
    cparams, wparams, wfields, preds = [], [], [], []
    for i, (T, C) in enumerate(test_draws):
        seed = 50000 + args.test_offset + i
        y = net.forward(features(n, T, C)[None])[0]
        preds.append(y)
        cparams.append(box_params(n, T, C, seed=seed))
        pw = box_params(n, T, C, seed=seed, init=INIT_FIELDS)
        wparams.append(pw)
        wfields.append(fields9_from_prediction(pw, y, prof_of(pw)))
    t2 = time.perf_counter()
    wres, _ = run_batch_to_steady(wparams, wfields, check_every=args.check_every, tol=args.tol,
                                  max_steps=args.max_steps, min_steps=0, tag="evaluation (warm):",
                                  consistent=args.consistent)
    cres, cout = run_batch_to_steady(cparams, check_every=args.check_every, tol=args.tol,
                                     max_steps=args.max_steps, min_steps=0, tag="evaluation (cold):")
    t_eval = time.perf_counter() - t2
    res, out, params = [], [], []
    for i in range(len(test_draws)):
        res += [cres[i], wres[i]]
        out += [cout[i], None]
        params += [cparams[i], wparams[i]]
    rows = []
    for i, (T, C) in enumerate(test_draws):
        c, w = res[2 * i], res[2 * i + 1]
        ok = bool(c.converged and w.converged and abs(w.nusselt / c.nusselt - 1.0) < 5e-3)
        err = None
        if c.converged:   # the network's distance to the cold run's steady state (canonical position)
            pc = params[2 * i]
            yt = canonicalize(targets_from_fields(pc, out[2 * i], prof_of(pc)))
            err = round(float(np.linalg.norm(preds[i] - yt) / np.linalg.norm(yt)), 5)
        rows.append({"T": round(T, 2), "C": round(C, 3), "cold_steps": c.steps, "warm_steps": w.steps,
                     "cold_Nu": c.nusselt, "warm_Nu": w.nusselt, "cold_failed": c.failed,
                     "warm_failed": w.failed, "accepted": ok, "prediction_error": err})
    acc = [r for r in rows if r["accepted"]]
    speed = [r["cold_steps"] / max(r["warm_steps"], 1) for r in acc]
    summary = {
        "grid": list(n), "train_states": int(ntr), "test_draws": len(test_draws),
        "generation_s": round(t_gen, 1), "training_s": round(t_train, 1), "evaluation_s": round(t_eval, 1),
        "generation_steps_median": int(np.median(gen_steps)) if gen_steps else None,
        "network": args.load_net or "trained in this run",
        "final_train_loss": losses[-1] if losses else None, "first_epoch_loss": losses[0] if losses else None,
        "T_range": list(args.T_range), "C_range": list(args.C_range), "epochs": args.epochs,
        "test_seed": args.test_seed, "test_offset": args.test_offset,
        "check_every": args.check_every, "tol": args.tol, "max_steps": args.max_steps,
        "gen_max_steps": args.gen_max_steps or args.max_steps,
        "width": args.width, "modes": list(args.modes), "consistent_iterations": args.consistent,
        "train_draws": len(train_draws),
        "accepted": len(acc), "acceptance_rate": round(len(acc) / len(rows), 3),
        "prediction_error_median": (float(np.median([r["prediction_error"] for r in rows
                                                     if r["prediction_error"] is not None]))
                                    if any(r["prediction_error"] is not None for r in rows) else None),
        "steps_ratio_median_accepted": round(float(np.median(speed)), 3) if speed else None,
        "steps_ratio_mean_accepted": round(float(np.mean(speed)), 3) if speed else None,
        "total_cold_steps": int(sum(r["cold_steps"] for r in rows)),
        "total_steps_with_fallback": int(sum(r["warm_steps"] if r["accepted"] else r["warm_steps"] + r["cold_steps"]
                                             for r in rows)),
        "draws": rows,
    }
    print(json.dumps(summary), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
