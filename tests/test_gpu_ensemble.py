"""Ensembles (independent members in one replayed HIP graph) and the tile
autotuner: results must equal the same runs made member by member, eagerly,
bit for bit (the graph records exactly the eager launch sequence, and a
tile shape changes only which threads handle which lines)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _box(n, seed, mach):
    from public_amd import ic
    q = ic.periodic_box_state(n, seed, mach)
    h = ic.periodic_dx(n)
    return q, h, ic.cfl_dt(q, h)


def _fwd_member(n, seed, mach, **kw):
    q, h, dt = _box(n, seed, mach)
    m = dict(n=(n,) * 3, dx=(h,) * 3, mu0=1 / 1600, T0=1 / (1.4 * mach * mach), dt=dt, **kw)
    return m, q


def _standalone_direct(m, q, nsteps):
    from public_amd import Cruna
    kw = dict(m)
    with Cruna(kw.pop("n"), kw.pop("dx"), **kw) as g:
        g.set_state(q)
        g.run_direct(nsteps)
        g.synchronize()
        return g.get_state()


def _standalone_adjoint(m, qs, Q, nt_end, nsteps):
    from public_amd import Cruna
    kw = dict(m)
    with Cruna(kw.pop("n"), kw.pop("dx"), **kw) as g:
        g.trajectory_alloc(Q.shape[-1])
        for nt in range(1, Q.shape[-1] + 1):
            g.trajectory_set(nt, Q[..., nt - 1])
        g.set_adjoint_state(qs)
        g.run_adjoint(nsteps, nt_end)
        g.synchronize()
        return g.get_adjoint_state()


def _trajectory(m, q, nsteps):
    kw = dict(m)
    cfg = O.Cfg(kw.pop("n"), kw.pop("dx"), **{k: v for k, v in kw.items()
                                              if k in ("mu0", "T0", "dt", "filter_freq", "afilter_freq")})
    Q = np.zeros(q.shape + (nsteps,), order="F")
    s = q
    for nt in range(1, nsteps + 1):
        s = O.run_direct(cfg, s, 1, nt0=nt - 1)
        Q[..., nt - 1] = s
    return Q


def test_ensemble_equals_member_runs():
    from public_amd import ADJOINT, FORWARD, FP32_MIXED, Ensemble
    specs = [_fwd_member(32, 0, 0.3), _fwd_member(48, 1, 0.2, precision=FP32_MIXED),
             _fwd_member(16, 2, 0.1), _fwd_member(32, 3, 0.25, afilter_freq=(1, 1, 1)),
             _fwd_member(24, 4, 0.15, afilter_freq=(1, 1, 1))]
    adj = [3, 4]
    nt_traj = 7
    Qs, qsT = {}, {}
    for i in adj:
        Qs[i] = _trajectory(specs[i][0], specs[i][1], nt_traj)
        qsT[i] = np.asfortranarray(np.random.default_rng(10 + i).standard_normal(specs[i][1].shape))
    with Ensemble([s[0] for s in specs]) as ens:
        for i, (m, q) in enumerate(specs):
            if i in adj:
                g = ens[i]
                g.trajectory_alloc(nt_traj)
                for nt in range(1, nt_traj + 1):
                    g.trajectory_set(nt, Qs[i][..., nt - 1])
                g.set_adjoint_state(qsT[i])
                ens.set_role(i, ADJOINT, nt_traj)
            else:
                ens[i].set_state(q)
                ens.set_role(i, FORWARD, 0)
        ens.run(4)   # one eager step, three graph replays
        ens.run(2)   # replays of the recorded graph
        out = [ens[i].get_adjoint_state() if i in adj else ens[i].get_state() for i in range(len(specs))]
    for i, (m, q) in enumerate(specs):
        ref = (_standalone_adjoint(m, qsT[i], Qs[i], nt_traj, 6) if i in adj
               else _standalone_direct(m, q, 6))
        assert np.array_equal(out[i], ref), (i, float(np.abs(out[i] - ref).max()))


def test_ensemble_with_filter_period():
    """Filter frequencies 2 and 3: the recorded period is 6 steps."""
    from public_amd import FORWARD, Ensemble
    specs = [_fwd_member(32, 5, 0.3, filter_freq=(2, 1, 3)), _fwd_member(16, 6, 0.2, filter_freq=(1, 2, 1))]
    with Ensemble([s[0] for s in specs]) as ens:
        for i, (m, q) in enumerate(specs):
            ens[i].set_state(q)
            ens.set_role(i, FORWARD, 0)
        ens.run(15)
        out = [ens[i].get_state() for i in range(2)]
    for i, (m, q) in enumerate(specs):
        assert np.array_equal(out[i], _standalone_direct(m, q, 15)), i


def test_autotune_keeps_results():
    from public_amd import Cruna
    m, q = _fwd_member(64, 7, 0.3)
    kw = dict(m)
    with Cruna(kw.pop("n"), kw.pop("dx"), **kw) as g:
        g.set_state(q)
        tl = g.autotune(steps=1)
        assert all(t >= 1 and (t & (t - 1)) == 0 for t in tl)
        assert np.array_equal(g.get_state(), q)  # the state was restored
        g.run_direct(3)
        g.synchronize()
        got = g.get_state()
    assert np.array_equal(got, _standalone_direct(m, q, 3))


def test_ensemble_rerecords_after_member_setters_and_runs_new_paths():
    """Graphs recorded by a first run are recorded again when a member changes
    through a setter (here a penalisation set on member 0 between runs: its
    buffers and kernels change); member 1 runs the complete-RHS boundary row
    (204) with its point-wise RK4 update and the weighted filter's copies
    inside the graph.  Both equal the same runs made eagerly, bit for bit."""
    from public_amd import FORWARD, Cruna, Ensemble
    m0, q0 = _fwd_member(32, 11, 0.3)
    m1, q1 = _fwd_member(32, 12, 0.2, wfilter_freq=(1, 1, 1), bc=[(204, 1, 32, 1, 1, 1, 32, 0, -1, 0)])
    n = 32
    x = np.arange(n) * m0["dx"][0]
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    phi = np.asfortranarray(0.7 + 0.2 * np.sin(X) * np.cos(Y) * np.sin(Z))
    with Ensemble([m0, m1]) as ens:
        ens[0].set_state(q0)
        ens[1].set_qref(q1)
        ens[1].set_state(q1)
        ens.set_role(0, FORWARD, 0)
        ens.set_role(1, FORWARD, 0)
        ens.run(4)
        ens[0].set_penalization(phi)
        ens.run(4)
        out = [ens[i].get_state() for i in range(2)]
    kw0 = dict(m0)
    with Cruna(kw0.pop("n"), kw0.pop("dx"), **kw0) as g:
        g.set_state(q0)
        g.run_direct(4)
        g.set_penalization(phi)
        g.run_direct(4, nt0=4)
        g.synchronize()
        ref0 = g.get_state()
    kw1 = dict(m1)
    with Cruna(kw1.pop("n"), kw1.pop("dx"), **kw1) as g:
        g.set_qref(q1)
        g.set_state(q1)
        g.run_direct(8)
        g.synchronize()
        ref1 = g.get_state()
    assert np.array_equal(out[0], ref0)
    assert np.array_equal(out[1], ref1)
