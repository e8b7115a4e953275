"""Pins of the oracle's deformed-grid path (grid_metric.f90 calc_metric and
ns_rho_rhou_p_grid_transform.f90 right_hand_side_direct_3d).

The composite routines cannot be compiled from the reference alone (they need
its MPI and HDF5 layers), so they are pinned by what the mathematics fixes:
  * on an undeformed grid the metric is the identity, the Jacobian 1, and the
    grid-transform RHS equals the Cartesian RHS;
  * on a y-stretched grid the metric is diag(y_eta, 1, y_eta) with
    J = 1 / y_eta, a uniform state has zero RHS (free-stream preservation of
    the conservative metric form), and the RHS converges at high order to its
    continuum limit, derived symbolically here from the reference's formulas
    (including the momentum viscous term, which the reference forms as
    J d/dxi_k (M(k,j) tau_ij) with tau built from the unscaled gradients:
    a(y) tau_i1,x + (a(y) tau_i2)_y + a(y) tau_i3,z for the y stretch, a = y_eta;
    and the heat flux whose outer derivatives take no h: dx_j (lambda T_j)_j).
"""
import numpy as np
import pytest
import sympy as sp

from oracle import oracle as O

LY, BETA = 8.0, 1.5


def grid(n, stretch=True):
    """x, z periodic on [0, 2pi); y open on [-LY/2, LY/2] with tanh clustering."""
    n1, n2, n3 = n
    dx = (2 * np.pi / n1, LY / (n2 - 1), 2 * np.pi / n3)
    x = np.arange(n1) * dx[0]
    yc = -LY / 2 + np.arange(n2) * dx[1]
    z = np.arange(n3) * dx[2]
    y = (LY / 2) * np.tanh(BETA * yc / (LY / 2)) / np.tanh(BETA) if stretch else yc.copy()
    Xc = np.stack(np.meshgrid(x, yc, z, indexing="ij"), axis=-1)
    X = np.stack(np.meshgrid(x, y, z, indexing="ij"), axis=-1)
    return dx, np.asfortranarray(X), np.asfortranarray(Xc), (x, y, z)


def cfg(n, dx, **kw):
    return O.Cfg(n, dx, scheme=(O.PERIODIC, O.OPEN, O.PERIODIC), **kw)


def test_metric_identity_on_undeformed_grid():
    n = (16, 24, 16)
    dx, X, Xc, _ = grid(n, stretch=False)
    Xm, Xj = O.calc_metric(cfg(n, dx), X, Xc)
    eye = np.eye(3)
    for k in range(3):
        for j in range(3):
            assert np.abs(Xm[..., k, j] - eye[k, j]).max() < 1e-13, (k, j)
    assert np.abs(Xj - 1).max() < 1e-13


def test_metric_of_y_stretch():
    n = (16, 48, 16)
    dx, X, Xc, (x, y, z) = grid(n)
    c = cfg(n, dx)
    Xm, Xj = O.calc_metric(c, X, Xc)
    # y_eta: the compact derivative of y(eta) with the open scheme
    Y = np.asfortranarray(X[..., 1])
    y_eta = O.diff3d(c, 1, Y, dx[1])
    for k in range(3):
        for j in range(3):
            if k != j:
                assert np.abs(Xm[..., k, j]).max() < 1e-12, (k, j)
    assert np.abs(Xm[..., 0, 0] - y_eta).max() < 1e-13
    assert np.abs(Xm[..., 1, 1] - 1).max() < 1e-13
    assert np.abs(Xm[..., 2, 2] - y_eta).max() < 1e-13
    assert np.abs(Xj * y_eta - 1).max() < 1e-13
    # and y_eta approximates the mapping's derivative
    yc = -LY / 2 + np.arange(n[1]) * dx[1]
    exact = BETA / np.tanh(BETA) / np.cosh(BETA * yc / (LY / 2)) ** 2
    assert np.abs(y_eta[0, :, 0] - exact).max() < 2e-3 * exact.max()


def _smooth_state(X):
    x, y, z = X[..., 0], X[..., 1], X[..., 2]
    q = np.zeros(X.shape[:3] + (5,), order="F")
    rho = 1 + 0.1 * np.sin(x) * np.cos(z) + 0.1 * np.exp(-y ** 2 / 4)
    u = 0.3 * np.tanh(y / 2) + 0.05 * np.sin(x + z)
    v = 0.05 * np.cos(x) * np.exp(-y ** 2 / 8)
    w = 0.05 * np.sin(z) * np.cos(0.3 * y)
    p = 1 / 1.4 + 0.05 * np.cos(x) * np.sin(z) * np.exp(-y ** 2 / 6)
    q[..., 0], q[..., 1], q[..., 2], q[..., 3], q[..., 4] = rho, rho * u, rho * v, rho * w, p
    return q


def test_gt_rhs_equals_cartesian_on_undeformed_grid():
    """dx = 1 in every direction so that the two heat-flux conventions agree."""
    n = (16, 17, 16)
    dx = (1.0, 1.0, 1.0)
    x = np.arange(n[0]) * 1.0
    yc = np.arange(n[1]) * 1.0 - 8.0
    z = np.arange(n[2]) * 1.0
    Xc = np.asfortranarray(np.stack(np.meshgrid(x, yc, z, indexing="ij"), axis=-1))
    s = 2 * np.pi / 16
    q = _smooth_state(np.stack([Xc[..., 0] * s, Xc[..., 1] / 2, Xc[..., 2] * s], axis=-1))
    c = cfg(n, dx, mu0=0.02, T0=1.0 / 1.4)
    Xm, Xj = O.calc_metric(c, Xc, Xc)
    ref = O.rhs_direct(c, q, heat_unit_spacing=1)
    got = O.rhs_direct_gt(c, q, Xm, Xj)
    for v in range(5):
        e = np.linalg.norm(got[..., v] - ref[..., v]) / np.linalg.norm(ref[..., v])
        assert e < 1e-12, (v, e)
    # the metric in the config selects the grid-transform RHS
    cm = cfg(n, dx, mu0=0.02, T0=1.0 / 1.4, metric=(Xm, Xj))
    assert np.array_equal(O.rhs_direct(cm, q), got)


def test_gt_free_stream_preservation():
    n = (16, 32, 16)
    dx, X, Xc, _ = grid(n)
    c = cfg(n, dx, mu0=0.01, T0=1.0 / 1.4)
    Xm, Xj = O.calc_metric(c, X, Xc)
    q = np.zeros(n + (5,), order="F")
    q[..., 0], q[..., 1], q[..., 2], q[..., 3], q[..., 4] = 1.2, 0.3, -0.2, 0.1, 0.9
    rhs = O.rhs_direct_gt(c, q, Xm, Xj)
    assert np.abs(rhs).max() < 1e-12


def _continuum_rhs(mu0, T0, gamma, dxs):
    """Sympy continuum limit of the reference grid-transform RHS on the y-stretched grid."""
    x, y, z = sp.symbols("x y z", real=True)
    rho = 1 + sp.Rational(1, 10) * sp.sin(x) * sp.cos(z) + sp.Rational(1, 10) * sp.exp(-y ** 2 / 4)
    u = sp.Rational(3, 10) * sp.tanh(y / 2) + sp.Rational(1, 20) * sp.sin(x + z)
    v = sp.Rational(1, 20) * sp.cos(x) * sp.exp(-y ** 2 / 8)
    w = sp.Rational(1, 20) * sp.sin(z) * sp.cos(sp.Rational(3, 10) * y)
    p = 1 / sp.Float(1.4) + sp.Rational(1, 20) * sp.cos(x) * sp.sin(z) * sp.exp(-y ** 2 / 6)
    X = (x, y, z)
    U = (u, v, w)
    # a(y) = dy/deta of y = (LY/2) tanh(BETA eta / (LY/2)) / tanh(BETA)
    eta = (LY / 2) / BETA * sp.atanh(y * sp.tanh(BETA) / (LY / 2))
    a = 1 / sp.diff(eta, y)
    Tm = p / rho  # R = W = 1
    mu = mu0 * (Tm / T0) ** sp.Rational(3, 2) * (T0 + sp.Float(110.4)) / (Tm + sp.Float(110.4))
    lam = mu * (gamma / (gamma - 1)) / sp.Float(0.71)
    grad = [[sp.diff(U[i], X[j]) for j in range(3)] for i in range(3)]
    div = grad[0][0] + grad[1][1] + grad[2][2]
    tau = [[mu * (grad[i][j] + grad[j][i]) - (2 * mu * div / 3 if i == j else 0) for j in range(3)]
           for i in range(3)]
    r = [None] * 5
    r[0] = -sum(sp.diff(rho * U[j], X[j]) for j in range(3))
    for i in range(3):
        conv = -sum(sp.diff(rho * U[i] * U[j], X[j]) for j in range(3)) - sp.diff(p, X[i])
        visc = a * sp.diff(tau[i][0], x) + sp.diff(a * tau[i][1], y) + a * sp.diff(tau[i][2], z)
        r[1 + i] = conv + visc
    heat = sum(dxs[j] * sp.diff(lam * sp.diff(Tm, X[j]), X[j]) for j in range(3))
    diss = sum(tau[i][j] * grad[i][j] for i in range(3) for j in range(3))
    adv = sum(U[j] * sp.diff(p, X[j]) for j in range(3))
    r[4] = (gamma - 1) * (-(gamma / (gamma - 1)) * sum(sp.diff(p * U[j], X[j]) for j in range(3))
                          + adv + heat + diss)
    return [sp.lambdify((x, y, z), e, "numpy") for e in r]


@pytest.mark.parametrize("mu0", [0.0, 0.05])
def test_gt_rhs_converges_to_continuum(mu0):
    gamma, T0 = 1.4, 1.0 / 1.4
    errs = []
    for n in [(16, 33, 16), (32, 65, 32)]:
        dx, X, Xc, _ = grid(n)
        c = cfg(n, dx, mu0=mu0, T0=T0)
        Xm, Xj = O.calc_metric(c, X, Xc)
        q = _smooth_state(X)
        got = O.rhs_direct_gt(c, q, Xm, Xj)
        fns = _continuum_rhs(mu0, T0, gamma, dx)
        e = []
        for vv in range(5):
            exact = np.broadcast_to(fns[vv](X[..., 0], X[..., 1], X[..., 2]), n)
            # interior in y: the open closures are lower order at the two faces
            sl = (slice(None), slice(4, -4), slice(None))
            e.append(np.linalg.norm((got[..., vv] - exact)[sl]) / np.linalg.norm(exact[sl]))
        errs.append(np.array(e))
    print("errors", errs)
    assert (errs[1] < 2e-3).all(), errs[1]
    assert (errs[0] / errs[1] > 12).all(), errs[0] / errs[1]
