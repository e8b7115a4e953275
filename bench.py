#!/usr/bin/env python3
"""Headline benchmark: RK4 grid-point updates per second on one MI355X.

Workload (the 256^3 periodic-box configuration): 3-D triply periodic
compressible Navier-Stokes on [0, 2pi)^3, gamma 1.4, Pr 0.71, Re 1600
(mu0 = 1/1600), power-law viscosity mu0 (T/T0)^0.76 with T0 the initial
temperature, seeded solenoidal random velocity E(k) ~ k^4 exp(-2 (k/4)^2) at
rms Mach 0.3, Lele 6th-order compact derivatives and the GV2000 compact
filter (alpha 0.35) every step in x, y and z, RK4 at CFL 0.5.  The state
lives in HBM for the whole timed region.

One JSON line is printed.  `value` is the FP64 throughput in Mpts/s
(N^3 * steps / s / 1e6); the FP32-mixed form (stencil/filter/RHS in FP32,
FP64 RK4 accumulator) is reported beside it with its error against FP64, and
the FP64 adjoint loop (over a forward trajectory recorded on the device) with
its own rate.
`roofline` describes the dominant kernel (largest share of the timed step):
its algorithmic bytes per launch (compulsory reads and writes of the sweep)
over its HIP-event-timed average duration, against 8 TB/s.  `cpu_baseline`
times one full step of a 128^3 box of the same workload on the host with the
reference's own compiled o6 cores (all host threads, and one thread), and
the restated oracle beside it.
`closure_64_vs_256` runs the committed learned closure at 64^3 against the
plain 256^3 FP64 solve over the same physical time from a held-out seed:
wall-clock speedup, velocity-spectrum error against the box-filtered fine
solve (with and without the closure), and the closure evaluation's
HIP-event time against the BF16 matrix-core peak.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0
# closure network: multiply-adds per point of its four 3x3x3 convolutions
# over the real channels (5 -> 32 -> 32 -> 32 -> 5)
CLOSURE_MACS_PER_POINT = 27 * (5 * 32 + 32 * 32 + 32 * 32 + 32 * 5)
# Algorithmic FP64-equivalent words per grid point and launch (compulsory
# reads + writes of each sweep; FP32-mixed fields count as half words).
SWEEP_WORDS = {
    "sweep1_x": 14.0,
    "sweep1_y": 20.0,
    "sweep1_z": 21.0,
    "sweep2_z": 12.0,
    "sweep2_y": 12.0,
    "sweep2_x_rk": 28.25,  # mean over the four stages: 27, 32, 32, 22
    # sweep1_z and sweep2_z fused (k_sweep12_z): sweep1_z's 21 words and the
    # two velocity gradients (u_x, v_y) that only the viscous part reads; the
    # other gradients, the viscosity and the momentum RHS pass between the
    # parts in registers and LDS
    "sweep12_z": 23.0,
    "filter_x": 10.0,
    "filter_y": 10.0,
    "filter_z": 10.0,
    # adjoint sweeps: q and qs read, rhs written (z) or read-modify-written;
    # the x sweep also carries the RK4 tail (30, 35, 35, 25 words)
    "asweep_z": 15.0,
    "asweep_y": 20.0,
    "asweep_x_rk": 31.25,
}
# words of each sweep that are stage-state / RHS / gradient fields held in the
# compute precision (the rest are FP64 state or accumulator words)
SWEEP_COMPUTE_WORDS = {
    "sweep1_x": 14.0, "sweep1_y": 20.0, "sweep1_z": 21.0, "sweep2_z": 12.0, "sweep2_y": 12.0,
    "sweep12_z": 23.0, "sweep2_x_rk": 15.75, "filter_x": 0.0, "filter_y": 0.0, "filter_z": 0.0,
    "asweep_z": 5.0, "asweep_y": 10.0, "asweep_x_rk": 12.5,
}
# The step-level byte model of the reference's dependency graph (4624 B per
# point per FP64 RK4 step incl. filtering; 2672 B in FP32-mixed).
MODEL_BYTES = {"f64": 4624.0, "fp32-mixed": 2672.0}
# General curvilinear grid (full metric per point): words per point and launch
# of the curvilinear sweeps.  x/y/z first-derivative sweeps: state 5 + metric
# row 3, six velocity-gradient, three temperature-gradient and five RHS
# fields written (x, also mu) or read-modify-written (y, z); viscous z/y:
# metric row 3 + J + mu + 3 temperature gradients + 6 velocity gradients, 4
# RHS fields read-modify-written; viscous x: the same 14 + 5 RHS read + the
# RK4 tail (mean of 15, 20, 20, 10 words over the stages).
CURV_WORDS = {"sweep1_x": 23.0, "sweep1_y": 36.0, "sweep1_z": 36.0, "sweep2_z": 22.0,
              "sweep2_y": 22.0, "sweep2_x_rk": 35.25, "filter_x": 10.0, "filter_y": 10.0,
              "filter_z": 10.0}
# per point and FP64 step: four stages of the six sweeps and three filter passes
CURV_MODEL_BYTES = 8.0 * (4 * (23 + 36 + 36 + 22 + 22 + 35.25) + 30)


def kernel_bytes_per_point(name, mixed):
    words = SWEEP_WORDS[name]
    if not mixed:
        return 8.0 * words
    cw = SWEEP_COMPUTE_WORDS[name]
    return 4.0 * cw + 8.0 * (words - cw)


def run_gpu(q0, n, h, dt, precision, steps, warmup, kw):
    from public_amd import Cruna
    g = Cruna((n,) * 3, (h,) * 3, precision=precision, dt=dt, filter_freq=(1, 1, 1), **kw)
    g.set_state(q0)
    g.run_direct(warmup)
    g.synchronize()
    g.profile(True)
    g.synchronize()
    t0 = time.perf_counter()
    g.run_direct(steps, nt0=warmup)
    g.synchronize()
    t1 = time.perf_counter()
    prof = g.profile_read()
    g.profile(False)
    q = g.get_state()
    g.close()
    return (t1 - t0), prof, q


def run_adjoint_gpu(q0, n, h, dt, steps, kw):
    """Forward run recording the trajectory on the device, then the timed
    adjoint loop over it (adjoint filters every step)."""
    from public_amd import Cruna
    g = Cruna((n,) * 3, (h,) * 3, dt=dt, filter_freq=(1, 1, 1), afilter_freq=(1, 1, 1), **kw)
    g.trajectory_alloc(steps)
    g.trajectory_record(True)
    g.set_state(q0)
    g.run_direct(steps)
    g.set_adjoint_state(np.asfortranarray(np.random.default_rng(1).standard_normal(q0.shape)))
    g.synchronize()
    g.profile(True)
    g.synchronize()
    t0 = time.perf_counter()
    g.run_adjoint(steps)
    g.synchronize()
    t1 = time.perf_counter()
    prof = g.profile_read()
    g.profile(False)
    qs = g.get_adjoint_state()
    g.close()
    return (t1 - t0), prof, bool(np.isfinite(qs).all())


def run_shear_layer(steps, warmup, n=(512, 256, 128), seed=0):
    """Temporal shear layer on a y-stretched grid, FP32-mixed: Lx, Ly, Lz = 60,
    40, 30, tanh clustering factor 2 around y = 0, u = 0.5 tanh(2y) with seeded
    white noise of amplitude 1e-3 windowed by exp(-y^2) on all three velocity
    components, Mach 0.5 on the velocity difference (c = 2), uniform T and p,
    Re 1000; periodic x/z, open y with characteristic rows (201) on both y
    faces and a sponge on the RHS (400, length 4 in y, amplitude 1)."""
    from public_amd import FP32_MIXED, FP64, OPEN, PERIODIC, SUTHERLAND, Cruna, ic
    from public_amd import geometry as geo
    L = (60.0, 40.0, 30.0)
    x0, x1 = (0.0, -L[1] / 2, 0.0), (L[0], L[1] / 2, L[2])
    scheme = (PERIODIC, OPEN, PERIODIC)
    Xc, dx = geo.cartesian_grid(n, x0, x1, scheme)
    X = geo.tanh_stretch_y(Xc, x0[1], x1[1], 2.0)
    rng = np.random.default_rng(seed)
    y = X[..., 1]
    win = np.exp(-y ** 2)
    gamma = 1.4
    q = np.zeros(n + (5,), order="F")
    q[..., 0] = 1.0
    q[..., 1] = 0.5 * np.tanh(2 * y) + 1e-3 * rng.standard_normal(n) * win
    q[..., 2] = 1e-3 * rng.standard_normal(n) * win
    q[..., 3] = 1e-3 * rng.standard_normal(n) * win
    q[..., 4] = 4.0 / gamma
    dt = geo.cfl_dt(q, X, 0.5, gamma)
    n1, n2, n3 = n
    bc = [(201, 1, n1, 1, 1, 1, n3, 0, -1, 0), (201, 1, n1, n2, n2, 1, n3, 0, 1, 0),
          (400, 1, n1, 1, n2, 1, n3)]
    kw = dict(scheme=scheme, gamma=gamma, R=1.0, W=1.0, Pr=0.71, mu0=1e-3, T0=4.0 / gamma,
              visc_law=SUTHERLAND, dt=dt, filter_freq=(1, 1, 1), bc=bc)
    with Cruna(n, dx, **kw) as g64:  # metric from FP64 compact derivatives
        M, J = geo.calc_metric(g64, X, Xc)
    sigma = geo.sponge_shape(X, x0, x1, (-1.0, 4.0, -1.0), 1.0)

    def make(precision):
        g = Cruna(n, dx, precision=precision, **kw)
        g.set_metric(M, J)
        g.set_qref(q)
        g.set_sponge_shape(sigma)
        g.set_state(q)
        return g

    g = make(FP32_MIXED)
    g.run_direct(warmup)
    g.synchronize()
    t0 = time.perf_counter()
    g.run_direct(steps, nt0=warmup)
    g.synchronize()
    t1 = time.perf_counter()
    qe = g.get_state()
    g.close()
    # the same run in FP64 (untimed) for the error of the mixed mode, with
    # the kinetic-energy history of both sampled every 10 steps
    total = warmup + steps
    hist = {}
    for prec in (FP64, FP32_MIXED):
        g = make(prec)
        ke, done = [ic.kinetic_energy(q)], 0
        while done < total:
            k = min(10, total - done)
            g.run_direct(k, nt0=done)
            done += k
            ke.append(ic.kinetic_energy(g.get_state()))
        hist[prec] = (np.array(ke), g.get_state())
        g.close()
    ref = hist[FP64][1]
    err = rel_errors(qe, ref, gamma)
    ke64, ke32 = hist[FP64][0], hist[FP32_MIXED][0]
    err["kinetic_energy_history"] = float(np.linalg.norm(ke32 - ke64) / np.linalg.norm(ke64))
    err["after_steps"] = total
    err["same_as_timed_run"] = bool(np.array_equal(hist[FP32_MIXED][1], qe))
    pts = float(n1) * n2 * n3
    return {"value": round(pts * steps / (t1 - t0) / 1e6, 2), "unit": "Mpts/s",
            "ms_per_step": round((t1 - t0) / steps * 1e3, 4), "steps": steps, "warmup": warmup,
            "grid": list(n), "dt": dt, "dtype": "fp32-mixed (f32 RHS, f64 state/RK/boundaries)",
            "finite": bool(np.isfinite(qe).all()),
            "max_abs_v": float(np.abs(qe[..., 2] / qe[..., 0]).max()),
            "err_vs_fp64": err}


def run_curvilinear(steps, warmup, n=256, seed=0, mach=0.3, amp=0.1):
    """The periodic-box workload (same state, viscosity law and filters) on a
    general curvilinear grid: the triply periodic wavy mesh
    x = xi + amp sin(eta) sin(zeta), y = eta + amp sin(xi) sin(zeta),
    z = zeta + amp sin(xi) sin(eta) on [0, 2pi)^3, metric from calc_metric's
    conservative form with the library's own derivatives; FP64, RK4 with
    the x, y, z filters every step."""
    from public_amd import POWER, Cruna, ic
    from public_amd import geometry as geo
    gamma = 1.4
    h = ic.periodic_dx(n)
    xi = np.arange(n) * h
    Xc = np.stack(np.meshgrid(xi, xi, xi, indexing="ij"), axis=-1)
    s, t, u = Xc[..., 0], Xc[..., 1], Xc[..., 2]
    X = np.stack([s + amp * np.sin(t) * np.sin(u), t + amp * np.sin(s) * np.sin(u),
                  u + amp * np.sin(s) * np.sin(t)], axis=-1)
    del s, t, u
    X, Xc = np.asfortranarray(X), np.asfortranarray(Xc)
    q0 = ic.periodic_box_state(n, seed, mach, gamma)
    dt = geo.cfl_dt(q0, X, 0.5, gamma)
    kw = dict(gamma=gamma, R=1.0, W=1.0, Pr=0.71, mu0=1.0 / 1600.0, T0=1.0 / (gamma * mach ** 2),
              visc_law=POWER, visc_exp=0.76, filter_alpha=0.35, dt=dt, filter_freq=(1, 1, 1))
    g = Cruna((n,) * 3, (h,) * 3, **kw)
    M, J = geo.calc_metric(g, X, Xc)
    del X, Xc
    g.set_metric(M, J)
    del M, J
    g.set_state(q0)
    g.run_direct(warmup)
    g.synchronize()
    g.profile(True)
    g.synchronize()
    t0 = time.perf_counter()
    g.run_direct(steps, nt0=warmup)
    g.synchronize()
    t1 = time.perf_counter()
    prof = g.profile_read()
    g.profile(False)
    qe = g.get_state()
    g.close()
    pts = float(n) ** 3
    ms = (t1 - t0) / steps * 1e3
    kernels = {k: {"avg_ms": round(v[0] / v[1], 4), "launches": v[1],
                   "GBps": round(8.0 * CURV_WORDS[k] * pts / (v[0] / v[1] / 1e3) / 1e9, 1)}
               for k, v in prof.items() if k in CURV_WORDS}
    return {"value": round(pts * steps / (t1 - t0) / 1e6, 2), "unit": "Mpts/s",
            "ms_per_step": round(ms, 4), "steps": steps, "warmup": warmup, "grid": [n, n, n],
            "dt": dt, "dtype": "f64", "mesh": f"triply periodic wavy mesh, amplitude {amp}",
            "model_bytes_per_point": CURV_MODEL_BYTES,
            "step_frac_of_hbm": round(CURV_MODEL_BYTES * pts / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "kernels": kernels, "finite": bool(np.isfinite(qe).all())}


def run_ensemble(steps, nmembers=32, seed=0, autotune=True, subset_steps=10):
    """The batched workload of public_amd.workloads: periodic boxes with N
    drawn from {48, 64, 96, 128}^3 and Ma from U[0.05, 0.3] (seeded); the
    first half forward RK4 runs, the second half adjoint runs backward over
    forward trajectories checkpointed every `subset_steps` steps (the
    checkpoints are written by an untimed forward run; the adjoint recomputes
    each subset from its checkpoint inside the timed region).  Timed: all
    members advanced by `steps` steps, (a) one member after the other,
    eagerly, (b) together through cruna_hip_ensemble_run's replayed member
    graphs, (c) through its batched launches with one launch per sweep kind
    over all members, and (d) through the default batched launches: four
    bins of members, each bin one launch per sweep kind over its members,
    the bins side by side.  Mpts/s counts N^3 per member and step (an
    adjoint step's recompute is not counted as work)."""
    from public_amd import ADJOINT, FORWARD, Cruna, Ensemble, workloads
    specs, states, adj = workloads.ensemble_members(nmembers, seed)
    sizes = [s["n"][0] for s in specs]
    pts = float(sum(int(s) ** 3 for s in sizes))
    subsets = steps // subset_steps

    def prepare(get_member):
        for i in range(nmembers):
            g = get_member(i)
            g.set_state(states[i])
            if i in adj:
                g.subsets_alloc(subsets, subset_steps)
                g.run_direct_subsets(subsets, subset_steps)
                g.set_adjoint_state(workloads.final_adjoint_state(states[i].shape, seed, i))
                g.synchronize()

    # (a) one after the other, eagerly
    solo = []
    for i in range(nmembers):
        kw = dict(specs[i])
        solo.append(Cruna(kw.pop("n"), kw.pop("dx"), **kw))
    prepare(lambda i: solo[i])
    t0 = time.perf_counter()
    for i in range(nmembers):
        if i in adj:
            solo[i].run_adjoint_subsets(subsets, subset_steps)
        else:
            solo[i].run_direct(steps)
        solo[i].synchronize()
    t_serial = time.perf_counter() - t0
    ref = [solo[i].get_adjoint_state() if i in adj else solo[i].get_state() for i in range(nmembers)]
    for g in solo:
        g.close()
    # (b) the ensemble: per-member graphs on member streams, then batched
    # launches: one per sweep kind over all members (one bin), and the
    # default, four bins of about equal points side by side
    res = {}
    tuned = None
    for batching in (0, 1, -1):
        with Ensemble(specs) as ens:
            ens.set_batching(batching)
            prepare(lambda i: ens[i])
            if autotune:
                tuned = ens.autotune(1)
            for i in range(nmembers):
                ens.set_role(i, ADJOINT if i in adj else FORWARD, steps if i in adj else 0)
            t0 = time.perf_counter()
            ens.run(steps)
            secs = time.perf_counter() - t0
            same = all(np.array_equal(ens[i].get_adjoint_state() if i in adj else ens[i].get_state(), ref[i])
                       for i in range(nmembers))
            res[batching] = {"value": round(pts * steps / secs / 1e6, 2), "unit": "Mpts/s",
                             "seconds": round(secs, 4), "speedup_vs_serial": round(t_serial / secs, 3),
                             "batched_steps": ens.batched_steps(), "identical_to_serial": bool(same)}
    res[0]["note"] = "one graph per member and period, replayed on the member streams; includes recording"
    res[1]["note"] = ("one launch per sweep kind over all members (work list of member tiles), one graph per "
                      "subset period; includes recording")
    res[-1]["note"] = ("the members in 4 bins of about equal points, each bin one launch per sweep kind over its "
                       "members, the bins' graphs replayed side by side on 4 streams; includes recording")
    # HBM bytes the run moves by the sweeps' word counts: a forward step of
    # every member (the adjoint members' recomputed trajectories included),
    # a backward step of each adjoint member, the recorded trajectory slot
    # (5 words written) and the subset's checkpoint restore (10 words per
    # subset)
    fwd_words = 4 * sum(SWEEP_WORDS[k] for k in ("sweep1_x", "sweep1_y", "sweep12_z", "sweep2_y",
                                                   "sweep2_x_rk")) + SWEEP_WORDS["filter_y"] + SWEEP_WORDS["filter_z"]
    adj_words = 4 * sum(SWEEP_WORDS[k] for k in ("asweep_z", "asweep_y", "asweep_x_rk")) + 3 * SWEEP_WORDS["filter_x"]
    pts_adj = float(sum(int(sizes[i]) ** 3 for i in adj))
    step_bytes = 8.0 * (pts * fwd_words + pts_adj * (adj_words + 5.0 + 10.0 / subset_steps))
    for r in res.values():
        r["hbm_TBps_by_word_counts"] = round(step_bytes * steps / r["seconds"] / 1e12, 3)
    return {"members": nmembers, "sizes": [int(s) for s in sizes], "forward": nmembers - len(adj),
            "adjoint": len(adj), "steps": steps, "checkpoint_every": subset_steps,
            "points_per_step": pts, "bytes_per_step_by_word_counts": step_bytes,
            "serial_eager": {"value": round(pts * steps / t_serial / 1e6, 2), "unit": "Mpts/s",
                             "seconds": round(t_serial, 4),
                             "hbm_TBps_by_word_counts": round(step_bytes * steps / t_serial / 1e12, 3)},
            "member_graphs": res[0], "batched_one_launch": res[1], "batched": res[-1],
            "speedup": res[-1]["speedup_vs_serial"], "tile_lines": tuned,
            "identical_to_serial": all(r["identical_to_serial"] for r in res.values())}


def run_closure(coarse_steps, seed=2001, mach=0.3):
    """Learned closure at 64^3 against the plain 256^3 solve over the same
    physical time (coarse_steps steps at 4x the fine step), from a held-out
    seed: plain FP64 256^3 solve (timed), 64^3 FP64 with the committed
    closure and without it (timed), velocity-spectrum errors against the
    box-filtered fine solve, and the closure evaluation's timing on the
    context stream (HIP events) against the BF16 matrix-core peak."""
    from public_amd import Cruna, ic
    from public_amd import closure_study as CS
    model = CS.ClosureModel(os.path.join(ROOT, "public_amd", "data", "closure_64.npz"))
    m, f = model.S, model.factor
    nf = m * f
    q0 = ic.bandlimited_box_state(nf, seed, mach)
    dt_f = CS.fine_dt(q0, nf)
    kw = CS.physics(mach)
    with Cruna((nf,) * 3, (ic.periodic_dx(nf),) * 3, dt=dt_f, **kw) as g:
        g.set_state(q0)
        qb0 = g.box_filter_state(f, rhs=False)[0]
        g.run_direct(2)
        g.synchronize()
        g.set_state(q0)
        t0 = time.perf_counter()
        g.run_direct(coarse_steps * f)
        g.synchronize()
        t_fine = time.perf_counter() - t0
        ref = g.box_filter_state(f, rhs=False)[0]
    out = {"seed": seed, "mach": mach, "coarse_grid": m, "fine_grid": nf, "coarse_steps": coarse_steps,
           "fine_steps": coarse_steps * f, "fine_fp64_seconds": round(t_fine, 4)}
    cl = model.closure()
    try:
        for name, net in (("no_closure", None), ("closure", cl)):
            with Cruna((m,) * 3, (ic.periodic_dx(m),) * 3, dt=dt_f * f, **kw) as g:
                if net is not None:
                    g.set_closure(net)
                g.set_state(qb0)
                g.run_direct(1)
                g.synchronize()
                g.set_state(qb0)
                t0 = time.perf_counter()
                g.run_direct(coarse_steps)
                g.synchronize()
                secs = time.perf_counter() - t0
                q = g.get_state()
                res = {"seconds": round(secs, 4), "ms_per_step": round(secs / coarse_steps * 1e3, 4),
                       "speedup_vs_fine_fp64": round(t_fine / secs, 2),
                       "spectrum_error": round(CS.spectrum_error(q, ref), 4),
                       "finite": bool(np.all(np.isfinite(q)))}
                if net is not None:
                    g.set_state(qb0)
                    g.profile(True)
                    g.run_direct(10)
                    g.synchronize()
                    ms, n = g.profile_read()["closure"]
                    g.profile(False)
                    avg = ms / n
                    achieved = 2.0 * CLOSURE_MACS_PER_POINT * m ** 3 / (avg / 1e3) / 1e12
                    res["evaluation_ms"] = round(avg, 4)
                    res["roofline"] = {"bound": "mfma", "achieved": round(achieved, 2),
                                       "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                                       "frac": round(achieved / MFMA_BF16_PEAK_TFLOPS, 4),
                                       "algorithmic_flops_per_evaluation": 2.0 * CLOSURE_MACS_PER_POINT * m ** 3}
                out[name] = res
    finally:
        cl.close()
    out["validated_error_at_horizon"] = model.validated_error(coarse_steps)
    out["closure_used_by_fallback_rule"] = model.usable(coarse_steps)
    out["validated_horizon_coarse_steps"] = model.horizon()
    out["gain"] = [float(x) for x in model.gain]
    return out


def ke_history(q0, n, h, dt, precision, steps, every, kw):
    from public_amd import Cruna, ic
    g = Cruna((n,) * 3, (h,) * 3, precision=precision, dt=dt, filter_freq=(1, 1, 1), **kw)
    g.set_state(q0)
    hist = [ic.kinetic_energy(q0)]
    done = 0
    while done < steps:
        k = min(every, steps - done)
        g.run_direct(k, nt0=done)
        done += k
        hist.append(ic.kinetic_energy(g.get_state()))
    g.close()
    return np.array(hist)


def rel_errors(q, ref, gamma):
    from public_amd import ic
    out = {}
    names = ["rho", "rho_u", "rho_v", "rho_w", "p"]
    for v, name in enumerate(names):
        out[name] = float(np.linalg.norm(q[..., v] - ref[..., v]) / np.linalg.norm(ref[..., v]))
    rE, rEr = ic.rho_E(q, gamma), ic.rho_E(ref, gamma)
    out["rho_E"] = float(np.linalg.norm(rE - rEr) / np.linalg.norm(rEr))
    # the momentum as one vector field: a transverse component that is small
    # (a shear layer's rho v, rho w) has a large error relative to itself
    # but not relative to the momentum
    out["momentum"] = float(np.linalg.norm(q[..., 1:4] - ref[..., 1:4]) / np.linalg.norm(ref[..., 1:4]))
    return out


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_step_child(n, mode, threads, seed, mach):
    """Child process: one full step (RK4 + filters i1, i2, i3) of the periodic
    box on the host, timed end to end; prints one JSON line."""
    from oracle import oracle as O
    from public_amd import ic
    h = ic.periodic_dx(n)
    q = ic.periodic_box_state(n, seed, mach)
    kw = dict(mu0=1.0 / 1600.0, T0=1.0 / (1.4 * mach * mach), visc_law=O.POWER, visc_exp=0.76,
              dt=ic.cfl_dt(q, h), filter_freq=(1, 1, 1))
    O.set_threads(threads)
    if mode == "reference" and not O.use_reference_cores(threads=threads if threads > 1 else 0):
        print(json.dumps({"error": "oracle/_ref is not built"}))
        return
    cfg = O.Cfg((n,) * 3, (h,) * 3, **kw)
    t0 = time.perf_counter()
    O.run_direct(cfg, q, 1)
    el = time.perf_counter() - t0
    print(json.dumps({"seconds": el, "mpts": n ** 3 / el / 1e6}))


def cpu_baseline(n_sample=128, seed=0, mach=0.3):
    """The CPU path on the host cores, one full step (RK4 with its four RHS
    evaluations, then the three filter passes) of a periodic box of
    n_sample^3 points of the same workload, timed end to end:
      * reference cores, 1 thread: the reference's own compiled o6 derivative
        and filter cores (oracle/_ref, the reference's Fortran built in
        place) for every line solve, with the oracle's point-wise work;
      * reference cores, all threads: the reference's OpenMP o6 derivative
        core (diff_pen_core_o6_i_c_p_omp, tnr_omp = all host threads; the
        reference's filter has no threaded core) and the oracle's threaded
        point-wise loops;
      * port, all threads: the oracle's restated cores, threaded per line.
    Each runs in a child process (OMP_WAIT_POLICY=passive: the two OpenMP
    runtimes of a mixed step do not spin against each other)."""
    import subprocess
    threads = len(os.sched_getaffinity(0))
    env = dict(os.environ, OMP_WAIT_POLICY="passive")

    def run(mode, th):
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-step-child", mode, str(th),
                            str(n_sample), str(seed), str(mach)], capture_output=True, text=True, env=env,
                           timeout=600)
        try:
            return json.loads(r.stdout.strip().splitlines()[-1])
        except (ValueError, IndexError):
            return {"error": (r.stderr or r.stdout)[-300:]}

    one = run("reference", 1)
    allr = run("reference", threads)
    port = run("port", threads)
    model = _cpu_model()
    if "mpts" not in allr:
        return {"value": port.get("mpts"), "unit": "Mpts/s", "cores": threads, "kind": "port",
                "cpu": model, "sample": f"one full step of the {n_sample}^3 box (oracle/_ref unavailable)"}
    return {
        "value": round(allr["mpts"], 4), "unit": "Mpts/s", "cores": threads, "kind": "reference",
        "cpu": model,
        "sample": (f"one full step (RK4 + filters i1, i2, i3) of the {n_sample}^3 periodic box of the "
                   f"same workload on {threads} host threads: the reference's compiled OpenMP o6 "
                   f"derivative core and its serial o6 filter core for every line solve, the oracle's "
                   f"threaded point-wise work; {allr['seconds']:.2f} s"),
        "single_core": {"value": round(one["mpts"], 4) if "mpts" in one else None, "unit": "Mpts/s",
                        "cores": 1, "kind": "reference",
                        "seconds": round(one["seconds"], 2) if "seconds" in one else None},
        "port_all_cores": {"value": round(port["mpts"], 4) if "mpts" in port else None, "unit": "Mpts/s",
                           "cores": threads, "kind": "port",
                           "seconds": round(port["seconds"], 2) if "seconds" in port else None},
    }


def read_traffic(name, n, precision_tag):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if present."""
    path = os.path.join(ROOT, "profiles", f"traffic_{n}_{precision_tag}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 2 and parts[0] == name:
            try:
                return float(parts[1])
            except ValueError:
                return None
    return None


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--mach", type=float, default=0.3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--ke-every", type=int, default=20)
    ap.add_argument("--adjoint-steps", type=int, default=20,
                    help="adjoint steps timed after a recorded forward run (0: skip)")
    ap.add_argument("--shear-steps", type=int, default=50,
                    help="timed steps of the stretched-grid shear layer (512x256x128, FP32-mixed; 0: skip)")
    ap.add_argument("--curvilinear-steps", type=int, default=20,
                    help="timed steps of the 256^3 box on the wavy curvilinear mesh (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=100,
                    help="steps of the 32-member forward/adjoint ensemble (0: skip)")
    ap.add_argument("--closure-steps", type=int, default=300,
                    help="coarse steps of the 64^3 closure leg against the plain 256^3 solve (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-errors", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run 3 device-to-device copies of 1 GiB (for PMC scaling)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the multi-GPU "
              "(z-slab, partitioned tridiagonal) design is documented but not built", file=sys.stderr)
        sys.exit(2)

    from public_amd import FP32_MIXED, FP64, POWER, ic

    n = args.n
    gamma = 1.4
    h = ic.periodic_dx(n)
    q0 = ic.periodic_box_state(n, args.seed, args.mach, gamma)
    dt = ic.cfl_dt(q0, h, gamma)
    T0 = 1.0 / (gamma * args.mach ** 2)  # initial temperature with R = W = 1
    kw = dict(gamma=gamma, R=1.0, W=1.0, Pr=0.71, mu0=1.0 / 1600.0, T0=T0, visc_law=POWER,
              visc_exp=0.76, filter_alpha=0.35)

    secs64, prof64, q64 = run_gpu(q0, n, h, dt, FP64, args.steps, args.warmup, kw)
    secs32, prof32, q32 = run_gpu(q0, n, h, dt, FP32_MIXED, args.steps, args.warmup, kw)
    adjoint = None
    if args.adjoint_steps > 0:
        k = args.adjoint_steps
        secs_a, prof_a, finite = run_adjoint_gpu(q0, n, h, dt, k, kw)
        adjoint = {"value": round(float(n) ** 3 * k / secs_a / 1e6, 2), "unit": "Mpts/s",
                   "steps": k, "ms_per_step": round(secs_a / k * 1e3, 4), "finite": finite,
                   "kernels": {name: {"avg_ms": round(v[0] / v[1], 4), "launches": v[1],
                                      "GBps": round(kernel_bytes_per_point(name, False) * float(n) ** 3
                                                    / (v[0] / v[1] / 1e3) / 1e9, 1)}
                               for name, v in prof_a.items()}}
    shear = run_shear_layer(args.shear_steps, 2) if args.shear_steps > 0 else None
    curv = run_curvilinear(args.curvilinear_steps, 2, n) if args.curvilinear_steps > 0 else None
    ensemble = run_ensemble(args.ensemble_steps) if args.ensemble_steps > 0 else None
    closure = run_closure(args.closure_steps) if args.closure_steps > 0 else None
    pts = float(n) ** 3
    mpts64 = pts * args.steps / secs64 / 1e6
    mpts32 = pts * args.steps / secs32 / 1e6

    # dominant kernel of the FP64 step
    dom = max(prof64.items(), key=lambda kv: kv[1][0])
    dname, (dms, dcount) = dom
    avg_s = dms / dcount / 1e3
    bytes_launch = kernel_bytes_per_point(dname, False) * pts
    achieved = bytes_launch / avg_s / 1e9
    traffic = read_traffic(dname, n, "fp64")
    roofline = {
        "kernel": dname,
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "algorithmic_bytes_per_launch": bytes_launch,
        "avg_launch_ms": round(avg_s * 1e3, 4),
    }
    kernels = {k: {"avg_ms": round(v[0] / v[1], 4), "launches": v[1],
                   "GBps": round(kernel_bytes_per_point(k, False) * pts / (v[0] / v[1] / 1e3) / 1e9, 1)}
               for k, v in prof64.items()}
    step_ms64 = secs64 / args.steps * 1e3
    step_ms32 = secs32 / args.steps * 1e3
    step_roofline = {
        "model_bytes_per_point": MODEL_BYTES["f64"],
        "frac": round(MODEL_BYTES["f64"] * pts / (step_ms64 / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "fp32_mixed_frac": round(MODEL_BYTES["fp32-mixed"] * pts / (step_ms32 / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
    }

    errors = None
    if not args.no_errors:
        errors = rel_errors(q32, q64, gamma)
        total = args.warmup + args.steps
        ke64 = ke_history(q0, n, h, dt, FP64, total, args.ke_every, kw)
        ke32 = ke_history(q0, n, h, dt, FP32_MIXED, total, args.ke_every, kw)
        errors["kinetic_energy_history"] = float(np.linalg.norm(ke32 - ke64) / np.linalg.norm(ke64))
        errors["after_steps"] = total

    if args.copy_calibration:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        nb = ctypes.c_size_t(1 << 30)
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        if hip.hipMalloc(ctypes.byref(a), nb) == 0 and hip.hipMalloc(ctypes.byref(b), nb) == 0:
            hip.hipMemset(a, 1, nb)
            for _ in range(3):
                hip.hipMemcpy(b, a, nb, 3)
            hip.hipDeviceSynchronize()
            hip.hipFree(a)
            hip.hipFree(b)

    cpu = None if args.no_cpu_baseline else cpu_baseline()

    out = {
        "metric": "grid-pt RK4 updates/s (Mpts/s) FP64 vs FP32-mixed; % HBM roofline; err vs FP64",
        "value": round(mpts64, 2),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms64, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded solenoidal random velocity field, no dataset)",
        "config": {"workload": f"periodic box {n}^3 compressible NS, RK4 + compact filter, Ma {args.mach}, Re 1600",
                   "grid": [n, n, n], "dt": dt, "seed": args.seed},
        "fp32_mixed": {"value": round(mpts32, 2), "unit": "Mpts/s", "ms_per_step": round(step_ms32, 4),
                       "speedup_vs_fp64": round(mpts32 / mpts64, 3), "err_vs_fp64": errors},
        "adjoint_fp64": adjoint,
        "shear_layer_fp32_mixed": shear,
        "curvilinear_fp64": curv,
        "ensemble_fp64": ensemble,
        "closure_64_vs_256": closure,
        "roofline": roofline,
        "step_roofline": step_roofline,
        "kernels_fp64": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "--cpu-step-child":
        _cpu_step_child(int(sys.argv[4]), sys.argv[2], int(sys.argv[3]), int(sys.argv[5]), float(sys.argv[6]))
    else:
        main()
