"""Learned subfilter closure for the periodic box: data, training, rollouts.

A coarse grid (64^3) with the learned closure stands in for a fine one
(256^3).  Everything runs through the library:

* data: seeded fine runs (the plain solver, `Cruna.run_direct`) are box
  filtered by the factor on the device (`Cruna.box_filter_state`), and the
  target of each filtered state q_bar is the subfilter residual of the RHS,
      target = box(RHS_fine(q_fine)) - RHS_coarse(q_bar),
  for the 5 state variables (rho, rho u, rho v, rho w, p);
* training: `Closure.train` (BF16 matrix-core forward and backward
  kernels, Adam) on batches of 8 windows of 32^3 points;
* rollouts: the coarse solver with the closure as its RHS force, from the
  box-filtered fine initial state, at the coarse CFL step (factor times the
  fine one), judged against the box-filtered plain fine solve by the
  velocity energy spectrum and the L2 error.  The plain fine solve is also
  the fallback: `ClosureModel.solve` uses the closure only when its
  validation error is within the threshold.

Run as a script it generates the data, trains, validates on held-out seeds
and writes the parameters and a summary:
    python -m public_amd.closure_study --out DIR
"""
import argparse
import json
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import ic
from .cruna import FP32_MIXED, FP64, POWER, Closure, Cruna

GAMMA = 1.4
P_REF = 1.0 / (GAMMA * 0.2 ** 2)  # mean pressure at Ma 0.2: the pressure output's reference


def physics(mach):
    """Periodic-box physics of the workload: Re 1600 on unit length and
    velocity, Pr 0.71, power-law viscosity about the initial temperature,
    compact filter every step.  The heat flux takes its derivatives with
    the grid spacing (heat_unit_spacing 0), so that the fine and the coarse
    grid solve the same equations (with unit spacing, as
    ns_rho_rhou_p.f90:189-197 forms it, the conduction term would scale
    with h^2 and differ 16-fold between 256^3 and 64^3)."""
    return dict(gamma=GAMMA, R=1.0, W=1.0, Pr=0.71, mu0=1.0 / 1600.0, T0=1.0 / (GAMMA * mach ** 2),
                visc_law=POWER, visc_exp=0.76, filter_alpha=0.35, filter_freq=(1, 1, 1), heat_unit_spacing=0)


def fine_dt(q, n):
    return ic.cfl_dt(q, ic.periodic_dx(n), GAMMA)


def box_average(q, f):
    """f^3 box average of (n, n, n, 5) Fortran-ordered fields (host)."""
    n = q.shape[0]
    m = n // f
    return np.asfortranarray(q.reshape(f, m, f, m, f, m, 5, order="F").mean(axis=(0, 2, 4)))


def velocity_spectrum(q):
    """E(k) = 1/2 sum over the shell |k| in [k - 1/2, k + 1/2) of |u_hat|^2 for
    the velocity (rho u / rho, ...) of a periodic state; k = 0 .. n/2."""
    n = q.shape[0]
    k1 = np.fft.fftfreq(n, d=1.0 / n)
    kx, ky, kz = np.meshgrid(k1, k1, k1, indexing="ij")
    shell = np.rint(np.sqrt(kx * kx + ky * ky + kz * kz)).astype(int)
    E = np.zeros(n // 2 + 1)
    for c in (1, 2, 3):
        uh = np.fft.fftn(q[..., c] / q[..., 0]) / n ** 3
        e = 0.5 * np.abs(uh) ** 2
        m = shell <= n // 2
        E += np.bincount(shell[m].ravel(), weights=e[m].ravel(), minlength=n // 2 + 1)
    return E


def spectrum_error(q, ref, kmax=None):
    """Relative L2 difference of the velocity spectra over shells 1..kmax
    (default: half the coarse grid's Nyquist wavenumber, the band the
    coarse grid resolves; beyond it the box filter's aliased tail and the
    coarse grid's own filter make the two incomparable)."""
    a, b = velocity_spectrum(q), velocity_spectrum(ref)
    kmax = kmax or (len(b) - 1) // 2
    return float(np.linalg.norm(a[1:kmax + 1] - b[1:kmax + 1]) / np.linalg.norm(b[1:kmax + 1]))


def state_error(q, ref):
    """Relative L2 error of rho, rho u (all components) and p."""
    out = {}
    for name, sl in (("rho", [0]), ("rho_u", [1, 2, 3]), ("p", [4])):
        a, b = q[..., sl], ref[..., sl]
        if name == "p":
            a, b = a - b.mean(), b - b.mean()
        out[name] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    return out


# ---------------------------------------------------------------------------
# data

def generate(closure, seeds, machs, n_fine=256, factor=4, steps=200, every=20, norm_runs=4, target="rhs",
             target_steps=1, log=print):
    """Training samples of seeded fine runs into `closure` (the first
    `norm_runs` runs also fix its normalisation).  Returns the statistics.

    target "rhs": the subfilter residual of the RHS,
        box(RHS_fine(q_f)) - RHS_coarse(box(q_f));
    target "step": the subfilter residual of `target_steps` coarse steps
    (RK4 and filter) per unit time,
        (box(q_f(t + K dt_c)) - Step_coarse^K(box(q_f(t)))) / (K dt_c),
    i.e. the RHS residual of the time-discrete coarse dynamics, which also
    carries what the coarse grid's filter and step size take away."""
    m = n_fine // factor
    hc = ic.periodic_dx(m)
    pending, held = [], []
    stats = None
    with ThreadPoolExecutor(max_workers=1) as pool:
        nxt = pool.submit(ic.bandlimited_box_state, n_fine, seeds[0], machs[0])
        for r, (seed, mach) in enumerate(zip(seeds, machs)):
            q0 = nxt.result()
            if r + 1 < len(seeds):
                nxt = pool.submit(ic.bandlimited_box_state, n_fine, seeds[r + 1], machs[r + 1])
            dt = fine_dt(q0, n_fine)
            t0 = time.time()
            kw = physics(mach)
            with Cruna((n_fine,) * 3, (ic.periodic_dx(n_fine),) * 3, dt=dt, precision=FP32_MIXED, **kw) as fine, \
                    Cruna((m,) * 3, (hc,) * 3, dt=dt * factor, **kw) as coarse:
                fine.set_state(q0)
                done = 0
                for s in range(0, steps + 1, every):
                    if s > done:
                        fine.run_direct(s - done, nt0=done)
                        done = s
                    if target == "rhs":
                        qb, rb = fine.box_filter_state(factor)
                        t = np.asfortranarray(rb - coarse.right_hand_side_direct_3d(qb))
                    else:
                        K = target_steps
                        if s + K * factor > steps:
                            break
                        qb, _ = fine.box_filter_state(factor, rhs=False)
                        fine.run_direct(K * factor, nt0=done)
                        done += K * factor
                        qb1, _ = fine.box_filter_state(factor, rhs=False)
                        coarse.set_state(qb)
                        coarse.run_direct(K, nt0=0)
                        t = np.asfortranarray((qb1 - coarse.get_state()) / (K * dt * factor))
                    pending.append((qb, t))
            if stats is None:
                held += pending
                pending = []
                if r + 1 == min(norm_runs, len(seeds)):
                    stats = normalisation(held)
                    closure.set_norm(stats["in_std"], stats["out_std"], P_REF)
                    for qb, t in held:
                        closure.add_sample(qb, t)
                    held = []
            else:
                for qb, t in pending:
                    closure.add_sample(qb, t)
                pending = []
            log(f"closure data: run {r + 1}/{len(seeds)} seed {seed} Ma {mach:.3f} dt {dt:.3e} "
                f"({time.time() - t0:.1f} s), samples {closure.num_samples()}")
    return stats


def normalisation(samples):
    """Input scales (fluctuations about each sample's means) and output
    scales (the pressure residual first divided by mean(p) / P_REF)."""
    ins = np.zeros(5)
    outs = np.zeros(5)
    for qb, t in samples:
        for c in range(5):
            ins[c] += np.var(qb[..., c] - qb[..., c].mean())
            scale = qb[..., 4].mean() / P_REF if c == 4 else 1.0
            outs[c] += np.mean((t[..., c] / scale) ** 2)
    k = len(samples)
    in_std = np.sqrt(ins / k)
    in_std[0] = max(in_std[0], 1e-6)
    return {"in_std": in_std.tolist(), "out_std": np.sqrt(outs / k).tolist(), "p_ref": P_REF}


# ---------------------------------------------------------------------------
# rollouts

def fine_reference(seed, mach, n_fine, factor, coarse_steps, every, precision=FP64, log=print):
    """Box-filtered plain fine solve at coarse steps 0, every, ... and the
    time the fine solver spent stepping (s)."""
    q0 = ic.bandlimited_box_state(n_fine, seed, mach)
    dt = fine_dt(q0, n_fine)
    out, secs = [], 0.0
    with Cruna((n_fine,) * 3, (ic.periodic_dx(n_fine),) * 3, dt=dt, precision=precision, **physics(mach)) as g:
        g.set_state(q0)
        out.append(g.box_filter_state(factor, rhs=False)[0])
        for s in range(every, coarse_steps + 1, every):
            t = time.perf_counter()
            g.run_direct(every * factor, nt0=(s - every) * factor)
            g.synchronize()
            secs += time.perf_counter() - t
            out.append(g.box_filter_state(factor, rhs=False)[0])
            log(f"fine reference: seed {seed} coarse step {s}/{coarse_steps}")
    return out, dt, secs


def coarse_rollout(q0, mach, dt, steps, every, closure=None, precision=FP64):
    """Coarse states at steps 0, every, ... and the stepping time (s)."""
    m = q0.shape[0]
    out, secs = [q0], 0.0
    with Cruna((m,) * 3, (ic.periodic_dx(m),) * 3, dt=dt, precision=precision, **physics(mach)) as g:
        if closure is not None:
            g.set_closure(closure)
        g.set_state(q0)
        g.run_direct(1, nt0=0)  # warm-up of the launch path, then restart
        g.synchronize()
        g.set_state(q0)
        for s in range(every, steps + 1, every):
            t = time.perf_counter()
            g.run_direct(every, nt0=s - every)
            g.synchronize()
            secs += time.perf_counter() - t
            out.append(g.get_state())
    return out, secs


def gain_label(g):
    """Result name of a rollout with inference gain g (a value or 5 values)."""
    g = np.broadcast_to(np.asarray(g, dtype=np.float64), (5,))
    if np.all(g == 1.0):
        return "closure"
    if np.all(g == g[0]):
        return f"closure_gain_{g[0]:g}"
    return "closure_gain_" + "_".join(f"{x:g}" for x in g)


def evaluate(closure, seed, mach, n_fine=256, factor=4, steps=300, every=50, log=print, precision=FP64,
             gains=(1.0,), fine_twin=False):
    """Held-out rollout without the closure and with it at each gain,
    against the filtered fine solve."""
    ref, dt_f, fine_secs = fine_reference(seed, mach, n_fine, factor, steps, every, log=log)
    dt_c = dt_f * factor
    res = {"seed": seed, "mach": mach, "coarse_steps": steps, "dt_coarse": dt_c, "fine_steps": steps * factor,
           "fine_secs": fine_secs, "spectrum_ref": velocity_spectrum(ref[-1]).tolist()}
    if fine_twin:
        # the same fine solve in FP32-mixed: its distance from the FP64 solve
        # is how far two solutions that differ only by rounding drift apart
        # (the chaotic growth of any difference), a floor for any model
        twin, _, _ = fine_reference(seed, mach, n_fine, factor, steps, every, precision=FP32_MIXED, log=log)
        res["fine_fp32_mixed"] = {"spectrum_error": [spectrum_error(q, r) for q, r in zip(twin, ref)],
                                  "state_error": [state_error(q, r) for q, r in zip(twin[1:], ref[1:])]}
        log(f"fine FP32-mixed vs FP64: final spectrum error {res['fine_fp32_mixed']['spectrum_error'][-1]:.4f}")
    runs = [("no_closure", None, 1.0)] + [(gain_label(g), closure, g) for g in gains if closure is not None]
    for name, cl, g in runs:
        if cl is not None:
            cl.set_gain(g)
        states, secs = coarse_rollout(ref[0], mach, dt_c, steps, every, cl, precision)
        res[name] = {
            "secs": secs,
            "spectrum_error": [spectrum_error(q, r) for q, r in zip(states, ref)],
            "state_error": [state_error(q, r) for q, r in zip(states[1:], ref[1:])],
            "finite": bool(np.all(np.isfinite(states[-1]))),
            "spectrum": velocity_spectrum(states[-1]).tolist(),
        }
        res[name]["speedup_vs_fine"] = fine_secs / secs
        log(f"rollout {name}: final spectrum error {res[name]['spectrum_error'][-1]:.4f}, "
            f"{secs:.3f} s vs fine {fine_secs:.3f} s")
    if closure is not None:
        closure.set_gain(1.0)
    return res


class ClosureModel:
    """A trained closure for one coarse grid with its inference gain and its
    validation: the largest spectrum error of the validation rollouts (one
    or several seeds) at coarse steps 0, every, 2 every, ...  `solve` runs
    the coarse grid with the closure when the validated error at the
    requested horizon is within margin x threshold, else the plain fine
    solver (the fallback).  The margin (default 1/2) allows for flows that
    drift faster than the validation seeds: a held-out Ma 0.3 seed reaches
    10.7 % at coarse step 125, where the worst of four validation seeds
    reads 8.6 %."""

    def __init__(self, path, threshold=0.10, margin=0.5):
        d = np.load(path)
        self.S = int(d["S"])
        self.factor = int(d["factor"])
        self.params = d["params"].astype(np.float32)
        self.in_std, self.out_std, self.p_ref = d["in_std"], d["out_std"], float(d["p_ref"])
        self.gain = np.broadcast_to(np.asarray(d["gain"], dtype=np.float64), (5,)).copy()
        self.val_steps = np.asarray(d["val_steps"], dtype=int)
        self.val_errors = np.asarray(d["val_errors"], dtype=float)
        self.threshold = threshold
        self.margin = margin

    def closure(self):
        c = Closure(self.S)
        c.set_norm(self.in_std, self.out_std, self.p_ref)
        c.set_params(self.params)
        c.set_gain(self.gain)
        return c

    def validated_error(self, coarse_steps):
        """Validated spectrum error at the first checkpoint at or beyond
        coarse_steps (inf beyond the validated horizon)."""
        k = np.searchsorted(self.val_steps, coarse_steps)
        return float(self.val_errors[k]) if k < len(self.val_steps) else np.inf

    def horizon(self):
        """Coarse steps from the start over which the validated error stays
        within margin x threshold (the last such checkpoint)."""
        ok = self.val_errors <= self.margin * self.threshold
        k = len(ok) if ok.all() else int(np.argmin(ok))
        return int(self.val_steps[k - 1]) if k > 0 else 0

    def usable(self, coarse_steps):
        """Whether the validated error stays within margin x threshold at
        every checkpoint up to the first one at or beyond coarse_steps."""
        k = int(np.searchsorted(self.val_steps, coarse_steps))
        if k >= len(self.val_steps):
            return False
        e = self.val_errors[:k + 1]
        return bool(np.all(np.isfinite(e)) and np.all(e <= self.margin * self.threshold))

    def solve(self, q_fine0, mach, fine_steps):
        """(box-filtered state after fine_steps fine steps' time, "closure" or
        "fine"): the coarse closure solution where validated, else the
        plain fine solution."""
        n = q_fine0.shape[0]
        dt = fine_dt(q_fine0, n)
        coarse_steps = fine_steps // self.factor
        if fine_steps % self.factor == 0 and n == self.S * self.factor and self.usable(coarse_steps):
            cl = self.closure()
            try:
                states, _ = coarse_rollout(box_average(q_fine0, self.factor), mach, dt * self.factor,
                                           coarse_steps, coarse_steps, cl)
            finally:
                cl.close()
            return states[-1], "closure"
        with Cruna((n,) * 3, (ic.periodic_dx(n),) * 3, dt=dt, **physics(mach)) as g:
            g.set_state(q_fine0)
            g.run_direct(fine_steps)
            return g.box_filter_state(self.factor, rhs=False)[0], "fine"


def save(path, closure, stats, factor, gain, val_steps, val_errors):
    np.savez(path, S=closure.S, factor=factor, params=closure.get_params(), in_std=np.array(stats["in_std"]),
             out_std=np.array(stats["out_std"]), p_ref=stats["p_ref"], gain=gain,
             val_steps=np.asarray(val_steps), val_errors=np.asarray(val_errors))


def revalidate(path, out_path, seeds, machs, n_fine=256, factor=4, steps=300, every=25, log=print):
    """Validation curve of a trained closure over several seeds: at each
    checkpoint the largest spectrum error among them, so that the closure's
    horizon holds for each of them.  Writes the closure file with that curve
    and returns each seed's curve."""
    model = ClosureModel(path)
    net = model.closure()
    curves = []
    try:
        for s, m in zip(seeds, machs):
            r = evaluate(net, s, m, n_fine, factor, steps, every=every, log=log, gains=[model.gain])
            curves.append(r[gain_label(model.gain)]["spectrum_error"])
    finally:
        net.close()
    d = dict(np.load(path))
    d["val_steps"] = np.arange(0, steps + 1, every)
    d["val_errors"] = np.max(np.array(curves), axis=0)
    d["val_seeds"] = np.asarray(seeds)
    d["val_machs"] = np.asarray(machs)
    np.savez(out_path, **d)
    return curves


def main():
    ap = argparse.ArgumentParser(description="generate closure data, train, validate")
    ap.add_argument("--out", required=True)
    ap.add_argument("--runs", type=int, default=64)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--every", type=int, default=20)
    ap.add_argument("--iters", type=int, default=16000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-decay-at", type=float, default=0.75,
                    help="fraction of the steps after which the learning rate drops tenfold")
    ap.add_argument("--val-seed", type=int, default=2000)
    ap.add_argument("--val-mach", type=float, default=0.25)
    ap.add_argument("--holdout", type=int, nargs="*", default=[2001, 2002, 2003])
    ap.add_argument("--holdout-mach", type=float, nargs="*", default=[0.3, 0.2, 0.15])
    ap.add_argument("--rollout-steps", type=int, default=300)
    ap.add_argument("--check-every", type=int, default=25)
    ap.add_argument("--n-fine", type=int, default=256)
    ap.add_argument("--factor", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--target", choices=("rhs", "step"), default="step")
    ap.add_argument("--target-steps", type=int, default=1, help="coarse steps of a 'step' target")
    ap.add_argument("--gains", nargs="*", default=["1", "1,1.1,1.1,1.1,1", "1,1.2,1.2,1.2,1", "1,1.3,1.3,1.3,1"],
                    help="inference gains to validate: a value or 5 comma-separated values each")
    ap.add_argument("--revalidate", default=None,
                    help="a trained closure file: only recompute its validation curve over --val-seeds")
    ap.add_argument("--val-seeds", type=int, nargs="*", default=[2000, 2004, 2005, 2006])
    ap.add_argument("--val-machs", type=float, nargs="*", default=[0.25, 0.3, 0.2, 0.1])
    args = ap.parse_args()
    os.makedirs(args.out, exist_ok=True)

    def log(msg):
        print(msg, flush=True)

    if args.revalidate:
        curves = revalidate(args.revalidate, os.path.join(args.out, "closure_params.npz"), args.val_seeds,
                            args.val_machs, args.n_fine, args.factor, args.rollout_steps, args.check_every, log)
        with open(os.path.join(args.out, "closure_revalidation.json"), "w") as f:
            json.dump({"seeds": args.val_seeds, "machs": args.val_machs, "every": args.check_every,
                       "spectrum_error": curves}, f, indent=1)
        log(json.dumps({"seeds": args.val_seeds, "final": [c[-1] for c in curves]}))
        return

    rng = np.random.default_rng(args.seed)
    seeds = list(range(1000, 1000 + args.runs))
    machs = rng.uniform(0.1, 0.3, size=args.runs).tolist()
    m = args.n_fine // args.factor
    net = Closure(m, seed=args.seed)
    t = time.time()
    stats = generate(net, seeds, machs, args.n_fine, args.factor, args.steps, args.every, target=args.target,
                     target_steps=args.target_steps, log=log)
    gen_secs = time.time() - t
    t = time.time()
    losses = []
    chunk = 500
    for i in range(0, args.iters, chunk):
        k = min(chunk, args.iters - i)
        lr = args.lr if i < args.lr_decay_at * args.iters else 0.1 * args.lr
        losses += net.train(k, lr=lr, batch=8, crop=32, seed=args.seed + i).tolist()
        log(f"closure training: step {i + k}/{args.iters} loss {np.mean(losses[-k:]):.4f}")
    train_secs = time.time() - t
    # the gain: the best final spectrum error on the validation seed
    gains = [[float(x) for x in s.split(",")] if "," in s else float(s) for s in args.gains]
    val = evaluate(net, args.val_seed, args.val_mach, args.n_fine, args.factor, args.rollout_steps,
                   every=args.check_every, log=log, gains=gains)
    best = None
    for g in gains:
        r = val[gain_label(g)]
        err = r["spectrum_error"][-1]
        if r["finite"] and np.isfinite(err) and (best is None or err < best[1]):
            best = (g, err, r["spectrum_error"])
    gain, _, val_errors = best if best else (0.0, np.inf, [np.inf] * (args.rollout_steps // args.check_every + 1))
    val_steps = list(range(0, args.rollout_steps + 1, args.check_every))
    log(f"closure gain {gain} (validation seed {args.val_seed})")
    evals = [evaluate(net, s, mm, args.n_fine, args.factor, args.rollout_steps, every=args.check_every, log=log,
                      gains=[gain], fine_twin=True) for s, mm in zip(args.holdout, args.holdout_mach)]
    gain = np.broadcast_to(np.asarray(gain, dtype=np.float64), (5,)).tolist()
    save(os.path.join(args.out, "closure_params.npz"), net, stats, args.factor, gain, val_steps, val_errors)
    summary = {"runs": args.runs, "steps": args.steps, "every": args.every, "target": args.target,
               "target_steps": args.target_steps,
               "samples": net.num_samples(), "data_secs": gen_secs, "train_secs": train_secs,
               "iters": args.iters, "loss_first": float(np.mean(losses[:50])), "loss_last": float(np.mean(losses[-50:])),
               "norm": stats, "gain": gain, "val_steps": val_steps, "validation": val, "holdout": evals}
    with open(os.path.join(args.out, "closure_summary.json"), "w") as f:
        json.dump(summary, f, indent=1)
    np.save(os.path.join(args.out, "closure_losses.npy"), np.array(losses, dtype=np.float32))
    log(json.dumps({k: v for k, v in summary.items() if k not in ("holdout", "validation")}))
    net.close()


if __name__ == "__main__":
    main()
