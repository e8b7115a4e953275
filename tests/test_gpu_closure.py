"""The learned closure's device kernels (BF16 matrix cores, FP32 accumulation)
against the numpy restatement with the same BF16 roundings (oracle/cnn.py):
the closure field, the loss and every layer's weight and bias gradients; the
closure as the additive force of the direct RHS; and training on the device
lowering the loss on a learnable target.

Tolerances: the device and the restatement sum in different orders in FP32
and FP64, so a value next to a BF16 rounding boundary can round to the
neighbouring BF16 value (relative 2^-8) in one of them; the relative L2
errors below allow for that (1e-2 for fields, 2e-2 for gradients).
"""
import numpy as np
import pytest

from oracle import cnn as CN

pytestmark = pytest.mark.gpu

S = 16
P_MEAN = 1.0 / (1.4 * 0.09)
IN_STD = np.array([0.02, 0.3, 0.3, 0.3, 0.05])
OUT_STD = np.array([0.5, 0.7, 0.7, 0.7, 0.9])
P_REF = 6.0


def _state(seed):
    """Periodic box velocity with density and pressure fluctuations."""
    from public_amd import ic
    q = ic.periodic_box_state(S, seed, 0.3)
    u, v, w = q[..., 1].copy(), q[..., 2].copy(), q[..., 3].copy()
    q[..., 0] = 1.0 + 0.02 * np.roll(u, 3, axis=1)
    q[..., 4] = P_MEAN + 0.05 * np.roll(w, 5, axis=0) * v
    q[..., 1:4] *= q[..., 0:1]
    return q


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _net(seed):
    from public_amd import Closure
    net = Closure(S, seed=seed)
    net.set_norm(IN_STD, OUT_STD, P_REF)
    return net


def test_forward_matches_oracle():
    net = _net(1)
    # larger weights than the initialisation so that every layer matters
    p = net.get_params() * 1.5
    net.set_params(p)
    q = _state(0)
    got = net.forward(q)
    ref = CN.closure(p.astype(np.float64), q, IN_STD, OUT_STD, P_REF)
    for v in range(5):
        e = _rel(got[..., v], ref[..., v])
        print("closure field", v, e)
        assert e < 1e-2, (v, e)
    net.close()


def test_gradients_match_oracle():
    net = _net(2)
    p = net.get_params()
    q = _state(1)
    rng = np.random.default_rng(0)
    target = np.asfortranarray(rng.standard_normal(q.shape) * OUT_STD)
    loss, g = net.gradients(q, target)
    x = CN.features(q, IN_STD)
    tn = CN.normalised_targets(q, target, OUT_STD, P_REF)
    rloss, rg = CN.loss_and_grads(p.astype(np.float64), x, tn)
    assert abs(loss - rloss) < 1e-3 * rloss, (loss, rloss)
    offs, n = CN.layout()
    for l, (w, b) in enumerate(offs):
        wsz = 27 * CN.CO[l] * CN.CI[l]
        ew = _rel(g[w:w + wsz], rg[w:w + wsz])
        eb = _rel(g[b:b + CN.CO[l]], rg[b:b + CN.CO[l]]) if np.any(rg[b:b + CN.CO[l]]) else 0.0
        print("layer", l, "weight grad", ew, "bias grad", eb)
        assert ew < 2e-2 and eb < 2e-2, (l, ew, eb)
    # padding stays inert: no gradient on padded input or output channels
    W0 = g[offs[0][0]:offs[0][0] + 27 * 32 * 16].reshape(27, 32, 16)
    assert not np.any(W0[:, :, 5:])
    W3 = g[offs[3][0]:offs[3][0] + 27 * 16 * 32].reshape(27, 16, 32)
    assert not np.any(W3[:, 5:, :])
    assert not np.any(g[offs[3][1] + 5:offs[3][1] + 16])
    net.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_closure_is_the_rhs_force(prec):
    from public_amd import FP32_MIXED, FP64, Cruna, ic
    net = _net(3)
    q = _state(2)
    h = ic.periodic_dx(S)
    kw = dict(mu0=1 / 1600, T0=1 / (1.4 * 0.09), dt=1e-3, precision=FP64 if prec == "fp64" else FP32_MIXED)
    with Cruna((S,) * 3, (h,) * 3, **kw) as g:
        plain = g.right_hand_side_direct_3d(q)
        g.set_closure(net)
        with_f = g.right_hand_side_direct_3d(q)
        g.set_closure(None)
        again = g.right_hand_side_direct_3d(q)
    f = net.forward(q)
    tol = 1e-13 if prec == "fp64" else 1e-6
    assert np.array_equal(plain, again)
    for v in range(5):
        scale = max(np.abs(plain[..., v]).max(), np.abs(f[..., v]).max())
        e = np.abs(with_f[..., v] - plain[..., v] - f[..., v]).max() / scale
        assert e < tol, (v, e)
    net.close()


def test_training_lowers_the_loss():
    """Targets made by a teacher network of the same shape are learnable."""
    teacher = _net(7)
    student = _net(8)
    qs = [_state(10 + s) for s in range(3)]
    for q in qs:
        student.add_sample(q, teacher.forward(q))
    losses = student.train(300, lr=2e-3, batch=4, crop=8, seed=1)
    print("loss", losses[0], losses[-1])
    assert np.all(np.isfinite(losses))
    assert losses[-20:].mean() < 0.2 * losses[:5].mean()
    teacher.close()
    student.close()


def test_window_loss_equals_whole_field_loss():
    """A window as large as the periodic sample, run with its halo of zeros
    outside, must see exactly the periodic field: its loss (before the
    first, zero-size, Adam step) equals the whole-field loss, wherever the
    window starts; and whole-sample batches (crop 0) give the same value."""
    net = _net(5)
    q = _state(4)
    rng = np.random.default_rng(1)
    target = np.asfortranarray(rng.standard_normal(q.shape) * OUT_STD)
    net.add_sample(q, target)
    whole, _ = net.gradients(q, target)
    for seed in range(3):
        w = net.train(1, lr=0.0, batch=1, crop=S, seed=seed)[0]
        assert abs(w - whole) < 1e-5 * whole, (seed, w, whole)
    full = net.train(1, lr=0.0, batch=2, crop=0, seed=0)[0]
    assert abs(full - whole) < 1e-5 * whole, (full, whole)
    net.close()


def test_box_filter_state():
    """Box averages of the device state and of its RHS against numpy means
    of the same state and of the RHS entry point's result."""
    from public_amd import Cruna, ic
    n, f = 32, 4
    q = ic.periodic_box_state(n, 3, 0.3)
    h = ic.periodic_dx(n)
    with Cruna((n,) * 3, (h,) * 3, mu0=1 / 1600, T0=1 / (1.4 * 0.09), dt=1e-3) as g:
        rhs = g.right_hand_side_direct_3d(q)
        g.set_state(q)
        qb, rb = g.box_filter_state(f)

    def box(a):
        m = n // f
        return a.reshape(f, m, f, m, f, m, 5, order="F").mean(axis=(0, 2, 4))

    assert np.abs(qb - box(q)).max() < 1e-13 * np.abs(q).max()
    assert np.abs(rb - box(rhs)).max() < 1e-13 * np.abs(rhs).max()


def test_attached_closure_cannot_be_freed():
    """A closure in use by a context refuses to be destroyed; detaching it
    (or destroying the context) releases it."""
    from public_amd import Cruna, CrunaError, ic
    net = _net(6)
    h = ic.periodic_dx(S)
    g = Cruna((S,) * 3, (h,) * 3)
    g.set_closure(net)
    with pytest.raises(CrunaError):
        net.close()
    g.set_closure(None)
    g.set_closure(net)
    g.close()
    net.close()


def test_closure_rejects_other_grids():
    from public_amd import OPEN, PERIODIC, Cruna, CrunaError
    net = _net(4)
    with Cruna((S, S, 2 * S), (0.1,) * 3) as g:
        with pytest.raises(CrunaError):
            g.set_closure(net)
    with Cruna((S,) * 3, (0.1,) * 3, scheme=(PERIODIC, OPEN, PERIODIC)) as g:
        with pytest.raises(CrunaError):
            g.set_closure(net)
    net.close()


def test_trained_closure_gains_over_plain_coarse_grid():
    """The committed closure (trained on the GPU by the library's own
    kernels on box-filtered 256^3 solves) against the same 64^3 solver
    without it, from the three held-out seeds (Ma 0.3, 0.2, 0.15) over 300
    coarse steps (1200 fine steps, the plain FP64 256^3 solve as the
    reference).  The 10 % bar on the filtered velocity-spectrum error holds
    up to the horizon the fallback rule validated (worst of four validation
    seeds within half the bar: 100 coarse steps) on every held-out seed, and
    at 300 steps, where the bar is out of reach (measured 25 / 21 / 22 %),
    the closure still cuts the plain coarse grid's error (58 / 39 / 27 %);
    the rule sends a 300-step request to the fine solver."""
    import os
    from public_amd import closure_study as CS
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    model = CS.ClosureModel(os.path.join(root, "public_amd", "data", "closure_64.npz"))
    h = model.horizon()
    assert h == 100, h
    assert model.usable(h) and not model.usable(300)
    net = model.closure()
    try:
        for seed, mach in ((2001, 0.3), (2002, 0.2), (2003, 0.15)):
            res = CS.evaluate(net, seed, mach, 256, 4, 300, every=h // 4, gains=[model.gain], log=print)
            errs = res[CS.gain_label(model.gain)]["spectrum_error"]
            plain = res["no_closure"]["spectrum_error"]
            print(f"seed {seed}: closure {errs[4]:.4f} at step {h}, {errs[-1]:.4f} at 300; plain {plain[-1]:.4f}")
            assert res[CS.gain_label(model.gain)]["finite"]
            assert errs[4] <= model.threshold, (seed, errs[4])
            assert errs[-1] < 0.85 * plain[-1], (seed, errs[-1], plain[-1])
    finally:
        net.close()
