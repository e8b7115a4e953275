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


# ---------------------------------------------------------------------------
# General curvilinear grid: every metric entry nonzero and varying along every
# direction.  x, z periodic on [0, 2pi) (each coordinate wraps along its own
# direction only, as calc_metric's periodic corrections assume), y open.
WAVY = (0.15, 0.2, 0.1, 0.05)


def _wavy_map(xi, eta, zeta, lib=np):
    a, b, c, e = WAVY
    X = xi + a * lib.sin(zeta) * lib.cos(0.5 * eta)
    Y = eta + b * lib.sin(xi) * lib.sin(zeta)
    Z = zeta + c * lib.sin(xi) + e * eta
    return X, Y, Z


def wavy_grid(n):
    n1, n2, n3 = n
    dx = (2 * np.pi / n1, LY / (n2 - 1), 2 * np.pi / n3)
    xi = np.arange(n1) * dx[0]
    eta = -LY / 2 + np.arange(n2) * dx[1]
    zeta = np.arange(n3) * dx[2]
    Xc = np.stack(np.meshgrid(xi, eta, zeta, indexing="ij"), axis=-1)
    X = np.stack(_wavy_map(Xc[..., 0], Xc[..., 1], Xc[..., 2]), axis=-1)
    return dx, np.asfortranarray(X), np.asfortranarray(Xc)


def _wavy_jacobian_matrix(Xc):
    """A[..., i, j] = dx_i / dxi_j of the wavy map, by hand."""
    a, b, c, e = WAVY
    xi, eta, zeta = Xc[..., 0], Xc[..., 1], Xc[..., 2]
    A = np.zeros(Xc.shape[:3] + (3, 3))
    A[..., 0, 0] = 1
    A[..., 0, 1] = -0.5 * a * np.sin(zeta) * np.sin(0.5 * eta)
    A[..., 0, 2] = a * np.cos(zeta) * np.cos(0.5 * eta)
    A[..., 1, 0] = b * np.cos(xi) * np.sin(zeta)
    A[..., 1, 1] = 1
    A[..., 1, 2] = b * np.sin(xi) * np.cos(zeta)
    A[..., 2, 0] = c * np.cos(xi)
    A[..., 2, 1] = e
    A[..., 2, 2] = 1
    return A


def test_curvilinear_metric_is_the_adjugate():
    """calc_metric on the wavy grid: X_metric(k,j) -> det(A) dxi_k/dx_j (the
    adjugate of A = dx/dxi, so that sum_k M(k,j) d/dxi_k = det(A) d/dx_j) and
    X_jacobian -> 1/det(A), at high order (a transposed or misplaced entry
    fails: every entry differs from its transpose here)."""
    errs = []
    for n in [(16, 33, 16), (32, 65, 32)]:
        dx, X, Xc = wavy_grid(n)
        Xm, Xj = O.calc_metric(cfg(n, dx), X, Xc)
        A = _wavy_jacobian_matrix(Xc)
        det = np.linalg.det(A)
        adj = det[..., None, None] * np.linalg.inv(A)
        sl = (slice(None), slice(4, -4), slice(None))
        e = max(np.abs(Xm[..., k, j] - adj[..., k, j])[sl].max() for k in range(3) for j in range(3))
        e = max(e, np.abs(Xj * det - 1)[sl].max())
        errs.append(e)
        # not symmetric: the test tells k from j
        assert np.abs(adj[..., 0, 1] - adj[..., 1, 0]).max() > 0.05
    assert errs[1] < 1e-6, errs
    assert errs[0] / errs[1] > 20, errs


def test_curvilinear_free_stream_preservation():
    """A uniform state has zero RHS on the wavy grid: the discrete metric
    identities of the conservative form hold to rounding in every direction."""
    n = (16, 32, 16)
    dx, X, Xc = wavy_grid(n)
    c = cfg(n, dx, mu0=0.01, T0=1.0 / 1.4)
    Xm, Xj = O.calc_metric(c, X, Xc)
    q = np.zeros(n + (5,), order="F")
    q[..., 0], q[..., 1], q[..., 2], q[..., 3], q[..., 4] = 1.2, 0.3, -0.2, 0.1, 0.9
    rhs = O.rhs_direct_gt(c, q, Xm, Xj)
    assert np.abs(rhs).max() < 1e-12


def _euler_physical(gamma):
    """The compressible Euler RHS in physical coordinates (no mapping in it)."""
    x, y, z = sp.symbols("x y z", real=True)
    rho, U, p = _smooth_fields(x, y, z)
    X = (x, y, z)
    r = [-sum(sp.diff(rho * U[j], X[j]) for j in range(3))]
    for i in range(3):
        r.append(-sum(sp.diff(rho * U[i] * U[j], X[j]) for j in range(3)) - sp.diff(p, X[i]))
    r.append((gamma - 1) * (-(gamma / (gamma - 1)) * sum(sp.diff(p * U[j], X[j]) for j in range(3))
                            + sum(U[j] * sp.diff(p, X[j]) for j in range(3))))
    return [sp.lambdify((x, y, z), e, "numpy") for e in r]


def _smooth_fields(x, y, z, lib=sp):
    """_smooth_state's fields as expressions."""
    rho = 1 + 0.1 * lib.sin(x) * lib.cos(z) + 0.1 * lib.exp(-y ** 2 / 4)
    u = 0.3 * lib.tanh(y / 2) + 0.05 * lib.sin(x + z)
    v = 0.05 * lib.cos(x) * lib.exp(-y ** 2 / 8)
    w = 0.05 * lib.sin(z) * lib.cos(0.3 * y)
    p = 1 / 1.4 + 0.05 * lib.cos(x) * lib.sin(z) * lib.exp(-y ** 2 / 6)
    return rho, (u, v, w), p


def _converges(got_exact_pairs, tol, ratio):
    errs = []
    for got, exact in got_exact_pairs:
        sl = (slice(None), slice(4, -4), slice(None))
        errs.append(np.array([np.linalg.norm((got[..., v] - exact[v])[sl]) / np.linalg.norm(exact[v][sl])
                              for v in range(5)]))
    print("errors", errs)
    assert (errs[1] < tol).all(), errs[1]
    assert (errs[0] / errs[1] > ratio).all(), errs[0] / errs[1]


def test_curvilinear_inviscid_rhs_converges_to_euler():
    """mu0 = 0: the grid-transform RHS on the wavy grid converges at high
    order to the Euler RHS evaluated in physical coordinates, which knows
    nothing of the mapping (an independent check of the contravariant
    fluxes, the metric identities and the metric placement)."""
    gamma = 1.4
    fns = _euler_physical(gamma)
    pairs = []
    for n in [(16, 33, 16), (32, 65, 32)]:
        dx, X, Xc = wavy_grid(n)
        c = cfg(n, dx, mu0=0.0, T0=1.0 / 1.4)
        Xm, Xj = O.calc_metric(c, X, Xc)
        q = _smooth_state(X)
        got = O.rhs_direct_gt(c, q, Xm, Xj)
        exact = [np.broadcast_to(f(X[..., 0], X[..., 1], X[..., 2]), n) for f in fns]
        pairs.append((got, exact))
    _converges(pairs, 2e-3, 12)


def _reference_form_in_xi(mu0, T0, gamma, dxs):
    """Continuum limit of the reference's grid-transform RHS written in the
    computational coordinates of the wavy map: fields are the physical ones
    composed with the map, X_metric = adj(dx/dxi), X_jacobian = 1/det, the
    derivatives exact; the heat flux's outer derivatives take no h
    (dxs[k] d/dxi_k) and the stresses are built from the gradients without
    Jacobian, as ns_rho_rhou_p_grid_transform.f90 forms them."""
    xi, eta, zeta = sp.symbols("xi eta zeta", real=True)
    S = (xi, eta, zeta)
    Xs = _wavy_map(xi, eta, zeta, lib=sp)
    A = sp.Matrix(3, 3, lambda i, j: sp.diff(Xs[i], S[j]))
    M = A.adjugate()  # M[k, j]
    J = 1 / A.det()
    rho, U, p = _smooth_fields(*Xs)
    Tm = p / rho
    mu = mu0 * (Tm / T0) ** sp.Rational(3, 2) * (T0 + sp.Float(110.4)) / (Tm + sp.Float(110.4))
    lam = mu * (gamma / (gamma - 1)) / sp.Float(0.71)

    def grad(f):  # physical gradient without Jacobian
        d = [sp.diff(f, s) for s in S]
        return [sum(M[k, j] * d[k] for k in range(3)) for j in range(3)]

    g = [grad(U[i]) for i in range(3)]
    pg, tg = grad(p), grad(Tm)
    uh = [sum(M[k, j] * U[j] for j in range(3)) for k in range(3)]
    div = g[0][0] + g[1][1] + g[2][2]
    tau = [[mu * (g[i][j] + g[j][i]) - (2 * mu * div / 3 if i == j else 0) for j in range(3)]
           for i in range(3)]
    r = [J * -sum(sp.diff(rho * uh[k], S[k]) for k in range(3))]
    for m in range(3):
        conv = -sum(sp.diff(rho * U[m] * uh[k], S[k]) for k in range(3)) - pg[m]
        visc = sum(sp.diff(sum(M[k, j] * tau[m][j] for j in range(3)), S[k]) for k in range(3))
        r.append(J * (conv + visc))
    heat = sum(M[k, j] * dxs[k] * sp.diff(lam * J * tg[j], S[k]) for j in range(3) for k in range(3))
    diss = J * sum(tau[i][j] * g[i][j] for i in range(3) for j in range(3))
    e = (gamma / (gamma - 1)) * -sum(sp.diff(p * uh[k], S[k]) for k in range(3))
    e = e + sum(U[j] * pg[j] for j in range(3)) + heat + diss
    r.append(J * (gamma - 1) * e)
    return [sp.lambdify((xi, eta, zeta), ex, "numpy") for ex in r]


def test_curvilinear_viscous_rhs_converges_to_reference_form():
    gamma, T0, mu0 = 1.4, 1.0 / 1.4, 0.05
    pairs = []
    for n in [(16, 33, 16), (32, 65, 32)]:
        dx, X, Xc = wavy_grid(n)
        fns = _reference_form_in_xi(mu0, T0, gamma, dx)
        c = cfg(n, dx, mu0=mu0, T0=T0)
        Xm, Xj = O.calc_metric(c, X, Xc)
        q = _smooth_state(X)
        got = O.rhs_direct_gt(c, q, Xm, Xj)
        exact = [np.broadcast_to(f(Xc[..., 0], Xc[..., 1], Xc[..., 2]), n) for f in fns]
        pairs.append((got, exact))
    _converges(pairs, 2e-3, 12)
