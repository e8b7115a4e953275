"""Pins of the oracle's Euler RHS (euler_rho_rhou_p.f90 right_hand_side_direct_3d):
  * it converges at high order to the compressible Euler equations on a
    periodic box (derived symbolically, in conservative form for mass and
    momentum and in the (gamma, u.grad p) form of the pressure equation);
  * it equals the NS RHS with mu0 = 0 to rounding (gamma D(p u) against
    gamma/(gamma-1) (gamma-1) D(p u)), which is how the device runs it."""
import numpy as np
import sympy as sp

from oracle import oracle as O


def _box(n):
    h = 2 * np.pi / n
    x = np.arange(n) * h
    return h, np.stack(np.meshgrid(x, x, x, indexing="ij"), axis=-1)


def _fields(x, y, z, lib=np):
    rho = 1 + 0.1 * lib.sin(x) * lib.cos(z) + 0.05 * lib.cos(y)
    u = (0.3 * lib.sin(y) + 0.05 * lib.sin(x + z), 0.05 * lib.cos(x) * lib.sin(z), 0.05 * lib.sin(z) * lib.cos(y))
    p = 1 / 1.4 + 0.05 * lib.cos(x) * lib.sin(z) * lib.cos(y)
    return rho, u, p


def _state(X):
    rho, u, p = _fields(X[..., 0], X[..., 1], X[..., 2])
    q = np.zeros(X.shape[:3] + (5,), order="F")
    q[..., 0] = rho
    for i in range(3):
        q[..., 1 + i] = rho * u[i]
    q[..., 4] = p
    return q


def test_euler_rhs_converges_to_euler_equations():
    gamma = 1.4
    x, y, z = sp.symbols("x y z", real=True)
    X = (x, y, z)
    rho, U, p = _fields(x, y, z, lib=sp)
    r = [-sum(sp.diff(rho * U[j], X[j]) for j in range(3))]
    for i in range(3):
        r.append(-sum(sp.diff(rho * U[i] * U[j], X[j]) for j in range(3)) - sp.diff(p, X[i]))
    r.append(-gamma * sum(sp.diff(p * U[j], X[j]) for j in range(3))
             + (gamma - 1) * sum(U[j] * sp.diff(p, X[j]) for j in range(3)))
    fns = [sp.lambdify((x, y, z), e, "numpy") for e in r]
    errs = []
    for n in (16, 32):
        h, G = _box(n)
        got = O.rhs_direct(O.Cfg((n,) * 3, (h,) * 3, equations=O.EQ_EULER), _state(G))
        errs.append([np.linalg.norm(got[..., v] - np.broadcast_to(fns[v](G[..., 0], G[..., 1], G[..., 2]), (n,) * 3))
                     / np.linalg.norm(np.broadcast_to(fns[v](G[..., 0], G[..., 1], G[..., 2]), (n,) * 3))
                     for v in range(5)])
    errs = np.array(errs)
    assert (errs[1] < 1e-4).all(), errs
    assert (errs[0] / errs[1] > 30).all(), errs


def test_euler_rhs_is_ns_without_viscosity():
    n = 24
    h, G = _box(n)
    q = _state(G)
    eu = O.rhs_direct(O.Cfg((n,) * 3, (h,) * 3, equations=O.EQ_EULER), q)
    ns = O.rhs_direct(O.Cfg((n,) * 3, (h,) * 3, mu0=0.0), q)
    for v in range(5):
        assert np.linalg.norm(eu[..., v] - ns[..., v]) <= 1e-13 * np.linalg.norm(eu[..., v]), v
    visc = O.rhs_direct(O.Cfg((n,) * 3, (h,) * 3, mu0=0.05), q)
    assert np.linalg.norm(visc - eu) > 1e-4 * np.linalg.norm(eu)


def test_euler_gt_rhs_pins():
    """The grid-transform Euler RHS (euler_rho_rhou_p_grid_transform.f90:33-140):
    on an undeformed grid (calc_metric of the computational grid: metric the
    identity, Jacobian 1) it equals the Cartesian Euler RHS; on a wavy
    deformed grid it equals the grid-transform NS RHS with mu0 = 0 (pinned
    by its own convergence tests) to rounding; a uniform stream stays at
    rest; a metric in the config selects it."""
    from tests.test_oracle_gt import _smooth_state, cfg, wavy_grid

    n = (16, 17, 16)
    dx = (2 * np.pi / 16, 0.5, 2 * np.pi / 16)
    xs = [np.arange(n[0]) * dx[0], -4 + np.arange(n[1]) * dx[1], np.arange(n[2]) * dx[2]]
    Xc = np.asfortranarray(np.stack(np.meshgrid(*xs, indexing="ij"), axis=-1))
    q = _smooth_state(Xc)
    c = cfg(n, dx, equations=O.EQ_EULER)
    Xm, Xj = O.calc_metric(c, Xc, Xc)
    ref = O.rhs_direct(c, q)
    got = O.rhs_direct_euler_gt(c, q, Xm, Xj)
    for v in range(5):
        e = np.linalg.norm(got[..., v] - ref[..., v]) / np.linalg.norm(ref[..., v])
        assert e < 1e-12, (v, e)

    n = (16, 24, 16)
    dx, X, Xc = wavy_grid(n)
    c = cfg(n, dx, equations=O.EQ_EULER)
    Xm, Xj = O.calc_metric(c, X, Xc)
    q = _smooth_state(X)
    got = O.rhs_direct_euler_gt(c, q, Xm, Xj)
    ns = O.rhs_direct_gt(cfg(n, dx, mu0=0.0), q, Xm, Xj)
    for v in range(5):
        e = np.linalg.norm(got[..., v] - ns[..., v]) / np.linalg.norm(ns[..., v])
        assert e < 1e-12, (v, e)
    assert np.array_equal(O.rhs_direct(cfg(n, dx, equations=O.EQ_EULER, metric=(Xm, Xj)), q), got)
    q0 = np.zeros(n + (5,), order="F")
    q0[..., 0], q0[..., 1], q0[..., 2], q0[..., 3], q0[..., 4] = 1.2, 0.3, -0.2, 0.1, 0.9
    assert np.abs(O.rhs_direct_euler_gt(c, q0, Xm, Xj)).max() < 1e-12


def test_euler_ev_phi_gt_rhs_pins():
    """The penalised grid-transform Euler RHS
    (euler_rho_rhou_p_ev_phi_grid_transform.f90:33-160): with a varying
    porosity on a wavy grid it equals the penalised grid-transform NS RHS
    with mu0 = 0 (pinned by convergence to the penalised Euler equations)
    to rounding; with porosity 1 it equals the grid-transform Euler RHS; the
    reference has no Cartesian penalised Euler module."""
    import pytest
    from tests.test_oracle_gt import _smooth_state, cfg, wavy_grid

    n = (16, 24, 16)
    dx, X, Xc = wavy_grid(n)
    Xm, Xj = O.calc_metric(cfg(n, dx), X, Xc)
    q = _smooth_state(X)
    phi = np.asfortranarray(0.7 + 0.2 * np.sin(X[..., 0]) * np.cos(0.4 * X[..., 1]) * np.sin(X[..., 2]))
    pen = (phi, 1.0 / phi)
    got = O.rhs_direct(cfg(n, dx, equations=O.EQ_EULER, metric=(Xm, Xj), penalization=pen), q)
    ns = O.rhs_direct(cfg(n, dx, mu0=0.0, metric=(Xm, Xj), penalization=pen), q)
    for v in range(5):
        e = np.linalg.norm(got[..., v] - ns[..., v]) / np.linalg.norm(ns[..., v])
        assert e < 1e-12, (v, e)
    one = np.ones(n, order="F")
    got1 = O.rhs_direct(cfg(n, dx, equations=O.EQ_EULER, metric=(Xm, Xj), penalization=(one, one)), q)
    ref1 = O.rhs_direct_euler_gt(cfg(n, dx, equations=O.EQ_EULER), q, Xm, Xj)
    for v in range(5):
        e = np.linalg.norm(got1[..., v] - ref1[..., v]) / np.linalg.norm(ref1[..., v])
        assert e < 1e-13, (v, e)
    with pytest.raises(ValueError, match="Cartesian"):
        cfg(n, dx, equations=O.EQ_EULER, penalization=pen)


def test_euler_adjoint_sponge_rows():
    """Type 400 in the Euler adjoint: set_boundary_condition_rhss gets the
    forward state (euler_rho_rhou_p.f90:384), so the row adds
    sponge_rhss_correct (sponge_layer_core.f90:107-151); the NS adjoint calls
    it without q, where the row is the empty sponge_rhss.  Pinned: shape 0
    gives the row-free adjoint RHS exactly, the added term is linear in the
    shape, and the NS adjoint ignores the row.  The matrix entries
    themselves (:131-145) are pinned against the reference's compiled
    routine in test_oracle_pins.py."""
    n = 16
    h, G = _box(n)
    q = _state(G)
    rng = np.random.default_rng(5)
    qs = np.asfortranarray(rng.standard_normal(q.shape))
    sig = np.asfortranarray(0.3 * rng.random((n,) * 3))
    zero = np.zeros((n,) * 3, order="F")
    row = [(400, 1, n, 1, n, 1, n)]
    eu = dict(equations=O.EQ_EULER)
    free = O.rhs_adjoint(O.Cfg((n,) * 3, (h,) * 3, **eu), qs, q)
    assert np.array_equal(O.rhs_adjoint(O.Cfg((n,) * 3, (h,) * 3, bc=row, **eu), qs, q, shape=zero), free)
    r1 = O.rhs_adjoint(O.Cfg((n,) * 3, (h,) * 3, bc=row, **eu), qs, q, shape=sig)
    r2 = O.rhs_adjoint(O.Cfg((n,) * 3, (h,) * 3, bc=row, **eu), qs, q, shape=np.asfortranarray(2 * sig))
    assert np.abs((r2 - free) - 2 * (r1 - free)).max() < 1e-12 * np.abs(r1 - free).max()
    assert np.abs(r1 - free).max() > 1e-3
    ns_free = O.rhs_adjoint(O.Cfg((n,) * 3, (h,) * 3), qs, q)
    assert np.array_equal(O.rhs_adjoint(O.Cfg((n,) * 3, (h,) * 3, bc=row), qs, q, shape=sig), ns_free)
