"""Pins of the oracle's volume-penalised direct RHS (ns_rho_rhou_p_ev_phi.f90
right_hand_side_direct_3d), which the reference only builds inside cases:
  * phi = 1 (the reference's 06_ns_ev_phi_pulse unit-test setting, whose
    volume_penalization_classic.f90 stores penalization(:,:,:,1:2) = 1) and
    any constant phi with 1/phi reduce it to the plain NS RHS with the
    unit-spacing heat flux;
  * with a varying porosity it converges at high order to the continuum
    equations of the penalised form, derived symbolically here:
      rho_t  = -(1/phi) div(phi rho u)
      m_i,t  = (1/phi) [-div(phi m_i u) + div(phi tau_i.)] - p_i
      p_t    = (gamma-1) [ (1/phi) (-gamma/(gamma-1) div(phi p u)
                                    + sum_k h_k^2 d_k(phi lambda d_k T))
                           + u.grad(p) + tau:grad(u) ]
    (h_k^2: both heat-flux derivatives of the reference take no h).
"""
import numpy as np
import pytest
import sympy as sp

from oracle import oracle as O


def _box(n):
    h = 2 * np.pi / n
    x = np.arange(n) * h
    X = np.stack(np.meshgrid(x, x, x, indexing="ij"), axis=-1)
    return h, X


def _fields(x, y, z, lib=np):
    rho = 1 + 0.1 * lib.sin(x) * lib.cos(z) + 0.05 * lib.cos(y)
    u = 0.3 * lib.sin(y) + 0.05 * lib.sin(x + z)
    v = 0.05 * lib.cos(x) * lib.sin(z)
    w = 0.05 * lib.sin(z) * lib.cos(y)
    p = 1 / 1.4 + 0.05 * lib.cos(x) * lib.sin(z) * lib.cos(y)
    phi = 0.7 + 0.2 * lib.sin(x) * lib.cos(y) * lib.sin(2 * z)
    return rho, (u, v, w), p, phi


def _state(X):
    rho, (u, v, w), p, phi = _fields(X[..., 0], X[..., 1], X[..., 2])
    q = np.zeros(X.shape[:3] + (5,), order="F")
    q[..., 0], q[..., 1], q[..., 2], q[..., 3], q[..., 4] = rho, rho * u, rho * v, rho * w, p
    return q, np.asfortranarray(phi)


@pytest.mark.parametrize("c", [1.0, 0.37])
def test_ev_phi_constant_porosity_is_plain_ns(c):
    n = 24
    h, X = _box(n)
    q, _ = _state(X)
    one = np.full((n,) * 3, c, order="F")
    kw = dict(mu0=0.02, T0=1.0 / 1.4)
    cfg = O.Cfg((n,) * 3, (h,) * 3, penalization=(one, 1.0 / one), **kw)
    got = O.rhs_direct(cfg, q)
    ref = O.rhs_direct(O.Cfg((n,) * 3, (h,) * 3, **kw), q, heat_unit_spacing=1)
    for v in range(5):
        e = np.linalg.norm(got[..., v] - ref[..., v]) / np.linalg.norm(ref[..., v])
        assert e < 1e-13, (v, e)


def _continuum(mu0, T0, gamma, h):
    x, y, z = sp.symbols("x y z", real=True)
    X = (x, y, z)
    rho, U, p, phi = _fields(x, y, z, lib=sp)
    Tm = p / rho  # R = W = 1
    mu = mu0 * (Tm / T0) ** sp.Rational(3, 2) * (T0 + sp.Float(110.4)) / (Tm + sp.Float(110.4))
    lam = mu * (gamma / (gamma - 1)) / sp.Float(0.71)
    grad = [[sp.diff(U[i], X[j]) for j in range(3)] for i in range(3)]
    div = grad[0][0] + grad[1][1] + grad[2][2]
    tau = [[mu * (grad[i][j] + grad[j][i]) - (2 * mu * div / 3 if i == j else 0) for j in range(3)]
           for i in range(3)]
    r = [-sum(sp.diff(phi * rho * U[j], X[j]) for j in range(3)) / phi]
    for i in range(3):
        flux = -sum(sp.diff(phi * rho * U[i] * U[j], X[j]) for j in range(3))
        visc = sum(sp.diff(phi * tau[i][j], X[j]) for j in range(3))
        r.append((flux + visc) / phi - sp.diff(p, X[i]))
    heat = sum(h * h * sp.diff(phi * lam * sp.diff(Tm, X[k]), X[k]) for k in range(3))
    conv = -(gamma / (gamma - 1)) * sum(sp.diff(phi * p * U[j], X[j]) for j in range(3))
    adv = sum(U[j] * sp.diff(p, X[j]) for j in range(3))
    diss = sum(tau[i][j] * grad[i][j] for i in range(3) for j in range(3))
    r.append((gamma - 1) * ((conv + heat) / phi + adv + diss))
    return [sp.lambdify((x, y, z), e, "numpy") for e in r]


def test_ev_phi_rhs_converges_to_penalised_equations():
    gamma, T0, mu0 = 1.4, 1.0 / 1.4, 0.05
    errs = []
    for n in (16, 32):
        h, X = _box(n)
        q, phi = _state(X)
        cfg = O.Cfg((n,) * 3, (h,) * 3, mu0=mu0, T0=T0, penalization=(phi, 1.0 / phi))
        got = O.rhs_direct(cfg, q)
        fns = _continuum(mu0, T0, gamma, h)
        e = []
        for v in range(5):
            exact = np.broadcast_to(fns[v](X[..., 0], X[..., 1], X[..., 2]), (n,) * 3)
            e.append(np.linalg.norm(got[..., v] - exact) / np.linalg.norm(exact))
        errs.append(np.array(e))
    print("errors", errs)
    assert (errs[1] < 1e-4).all(), errs[1]
    assert (errs[0] / errs[1] > 30).all(), errs[0] / errs[1]
    # the porosity matters: the plain RHS is far from the penalised limit
    n = 32
    h, X = _box(n)
    q, phi = _state(X)
    plain = O.rhs_direct(O.Cfg((n,) * 3, (h,) * 3, mu0=mu0, T0=T0), q, heat_unit_spacing=1)
    ev = O.rhs_direct(O.Cfg((n,) * 3, (h,) * 3, mu0=mu0, T0=T0, penalization=(phi, 1.0 / phi)), q)
    assert np.linalg.norm(plain - ev) / np.linalg.norm(ev) > 1e-2


# weighted filter (filter_weighted_iN with the o2_e_c_p core)
@pytest.mark.parametrize("d", [0, 1, 2])
def test_weighted_filter_transfer_function(d):
    """weight = strength = 1: q - F(F(q)) with F the 3-point stencil
    (q+ - 2 q + q-)/4, so a sine wave of step angle w is damped by
    1 - sin(w/2)^4 (the Nyquist mode is removed, constants are kept)."""
    shape = (16, 24, 32)
    cfg = O.Cfg(shape, (1.0, 1.0, 1.0))
    n = shape[d]
    for k in (1, 3, n // 4, n // 2):
        w = 2 * np.pi * k / n
        i = np.arange(n) * w + 0.3
        line = np.cos(i) if k != n // 2 else np.cos(np.arange(n) * np.pi)
        q = np.broadcast_to(line.reshape([-1 if a == d else 1 for a in range(3)]), shape).copy(order="F")
        got = O.filter_weighted4d(cfg, d, q)
        np.testing.assert_allclose(got, (1 - np.sin(w / 2) ** 4) * q, atol=1e-14)
    one = np.full(shape, 2.5, order="F")
    assert np.abs(O.filter_weighted4d(cfg, d, one) - one).max() < 1e-14


@pytest.mark.parametrize("d", [0, 1, 2])
def test_weighted_filter_conserves_the_weighted_sum(d):
    """sum(weight q) is kept along every line (F sums to zero over a periodic
    line, a0 + 2 a1 = 0: the pressure-tight property of the filter), while q
    changes; strength 0 leaves q unchanged."""
    shape = (16, 24, 32)
    rng = np.random.default_rng(3 + d)
    q = np.asfortranarray(rng.standard_normal(shape + (2,)))
    w = np.asfortranarray(rng.uniform(0.2, 1.0, shape))
    s = np.asfortranarray(rng.uniform(0.0, 1.0, shape))
    cfg = O.Cfg(shape, (1.0, 1.0, 1.0))
    got = O.filter_weighted4d(cfg, d, q, w, s)
    for v in range(2):
        before = (w * q[..., v]).sum(axis=d)
        after = (w * got[..., v]).sum(axis=d)
        assert np.abs(after - before).max() < 1e-12 * np.abs(w * q[..., v]).sum(axis=d).max()
    assert np.abs(got - q).max() > 1e-2
    zero = np.zeros(shape, order="F")
    assert np.array_equal(O.filter_weighted4d(cfg, d, q, w, zero), q)


# penalised deformed-grid RHS (ns_rho_rhou_p_ev_phi_grid_transform)
def test_ev_phi_gt_with_unit_porosity_is_the_grid_transform_rhs():
    from test_oracle_gt import _smooth_state, cfg as gt_cfg, wavy_grid
    n = (16, 24, 16)
    dx, X, Xc = wavy_grid(n)
    c0 = gt_cfg(n, dx, mu0=0.02, T0=1.0 / 1.4)
    Xm, Xj = O.calc_metric(c0, X, Xc)
    q = _smooth_state(X)
    one = np.ones(n, order="F")
    ref = O.rhs_direct_gt(c0, q, Xm, Xj)
    got = O.rhs_direct(gt_cfg(n, dx, mu0=0.02, T0=1.0 / 1.4, metric=(Xm, Xj), penalization=(one, one)), q)
    for v in range(5):
        e = np.linalg.norm(got[..., v] - ref[..., v]) / np.linalg.norm(ref[..., v])
        assert e < 1e-13, (v, e)


def _penalised_euler_physical(gamma):
    x, y, z = sp.symbols("x y z", real=True)
    X = (x, y, z)
    from test_oracle_gt import _smooth_fields
    rho, U, p = _smooth_fields(x, y, z)
    phi = 0.7 + 0.2 * sp.sin(x) * sp.cos(0.4 * y) * sp.sin(z)
    r = [-sum(sp.diff(phi * rho * U[j], X[j]) for j in range(3)) / phi]
    for i in range(3):
        r.append(-sum(sp.diff(phi * rho * U[i] * U[j], X[j]) for j in range(3)) / phi - sp.diff(p, X[i]))
    r.append((gamma - 1) * (-(gamma / (gamma - 1)) * sum(sp.diff(phi * p * U[j], X[j]) for j in range(3)) / phi
                            + sum(U[j] * sp.diff(p, X[j]) for j in range(3))))
    return [sp.lambdify((x, y, z), e, "numpy") for e in r], sp.lambdify((x, y, z), phi, "numpy")


def test_ev_phi_gt_inviscid_converges_to_penalised_euler():
    """mu0 = 0 on the wavy grid with a varying porosity: the penalised
    grid-transform RHS converges at high order to the penalised Euler
    equations in physical coordinates (rho_t = -(1/phi) div(phi rho u), ...),
    which know nothing of the mapping."""
    from test_oracle_gt import _converges, _smooth_state, cfg as gt_cfg, wavy_grid
    fns, phi_fn = _penalised_euler_physical(1.4)
    pairs = []
    for n in [(16, 33, 16), (32, 65, 32)]:
        dx, X, Xc = wavy_grid(n)
        c0 = gt_cfg(n, dx, mu0=0.0, T0=1.0 / 1.4)
        Xm, Xj = O.calc_metric(c0, X, Xc)
        phi = np.asfortranarray(np.broadcast_to(phi_fn(X[..., 0], X[..., 1], X[..., 2]), n))
        q = _smooth_state(X)
        got = O.rhs_direct(gt_cfg(n, dx, mu0=0.0, T0=1.0 / 1.4, metric=(Xm, Xj),
                                  penalization=(phi, 1.0 / phi)), q)
        exact = [np.broadcast_to(f(X[..., 0], X[..., 1], X[..., 2]), n) for f in fns]
        pairs.append((got, exact))
    _converges(pairs, 2e-3, 12)
