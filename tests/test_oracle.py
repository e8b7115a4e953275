"""Pins of the CPU oracle (oracle/cruna_oracle.c) -- CPU only.

Each function of the oracle is checked against something that does not
retype its formula:
  * LAPACK dgttrf/dgttrs            -> SciPy's LAPACK (same routines)
  * derivative / filter cores        -> the reference's own Fortran cores
                                        (oracle/_ref) and the committed golden
                                        vectors made from them; closed forms
  * closures, characteristic BCs,    -> the reference's own Fortran modules
    sponges
  * composite direct RHS             -> spectral evaluation of the same PDE,
                                        6th-order convergence; invariants
  * adjoint RHS                      -> duality with the linearised direct RHS
  * RK4                              -> closed-form amplification factor
"""
import os

import numpy as np
import pytest
from scipy.linalg import lapack

from oracle import oracle as O
from tests import refproc

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "cores.npz")
need_ref = pytest.mark.skipif(not refproc.available(), reason="oracle/_ref not built")


# --------------------------------------------------------------------------- LAPACK
@pytest.mark.parametrize("n", [3, 5, 16, 33])
@pytest.mark.parametrize("pivot", [False, True])
def test_dgttrf_dgttrs_vs_scipy(n, pivot):
    rng = np.random.default_rng(n + 100 * pivot)
    dl = rng.standard_normal(n - 1)
    d = rng.standard_normal(n) + (0.0 if pivot else 4.0)
    du = rng.standard_normal(n - 1)
    b = np.asfortranarray(rng.standard_normal((n, 3)))
    dl_s, d_s, du_s, du2_s, ipiv_s, info = lapack.dgttrf(dl, d, du)
    assert info == 0
    x_s, info = lapack.dgttrs(dl_s, d_s, du_s, du2_s, ipiv_s, b)
    a = [np.asfortranarray(v.copy()) for v in (dl, d, du)]
    du2 = np.zeros(n, order="F")
    ipiv = np.zeros(n, dtype=np.int32, order="F")
    assert O.lib().orc_dgttrf(n, a[0], a[1], a[2], du2, ipiv) == 0
    np.testing.assert_array_equal(ipiv, ipiv_s)
    np.testing.assert_allclose(a[1], d_s, rtol=1e-15, atol=0)
    x = b.copy(order="F")
    O.lib().orc_dgttrs(n, 3, a[0], a[1], a[2], du2, ipiv, x, n)
    np.testing.assert_allclose(x, x_s, rtol=1e-13, atol=1e-13)


# --------------------------------------------------------------------------- cores
def _rand_lines(n, nl, seed):
    return np.asfortranarray(np.random.default_rng(seed).standard_normal((n, nl)))


@need_ref
@pytest.mark.parametrize("scheme", [O.PERIODIC, O.OPEN])
@pytest.mark.parametrize("kind", ["diff", "filter"])
@pytest.mark.parametrize("n", [8, 24, 64])
def test_cores_vs_reference_fortran(scheme, kind, n):
    q = _rand_lines(n, 5, n)
    arg = 0.17 if kind == "diff" else 0.35
    ref = refproc.run("core", kind, scheme, q, arg)
    mine = O.diff_core(scheme, q, arg) if kind == "diff" else O.filter_core(scheme, q, arg)
    np.testing.assert_allclose(mine, ref, rtol=0, atol=1e-13 * np.abs(ref).max())


def test_cores_vs_golden():
    g = np.load(GOLDEN)
    for key in g.files:
        if not key.endswith("_in"):
            continue
        base = key[:-3]
        kind, scheme, n = base.split("_")[0], int(base.split("_")[1]), int(base.split("_")[2])
        q = g[key]
        arg = float(g[base + "_arg"])
        if kind == "filter":
            mine = O.filter_core(scheme, q, arg)
        else:
            dk = {"diffo4": O.DIFF_O4_EXPLICIT, "diffpade": O.DIFF_O4_PADE,
                  "diffo2": O.DIFF_O2_EXPLICIT}.get(kind, O.DIFF_O6_COMPACT)
            mine = O.diff_core(scheme, q, arg, dk)
        ref = g[base + "_out"]
        np.testing.assert_allclose(mine, ref, rtol=0, atol=1e-13 * np.abs(ref).max(), err_msg=base)


@need_ref
@pytest.mark.parametrize("scheme", [O.PERIODIC, O.OPEN])
@pytest.mark.parametrize("n", [8, 24, 64])
def test_o4_explicit_cores_vs_reference_fortran(scheme, n):
    """diff_pen_core_o4_e_c_{p,o} restated in the oracle against the reference's own."""
    q = _rand_lines(n, 5, 3 * n)
    ref = refproc.run("core", "diff_o4", scheme, q, 0.17)
    mine = O.diff_core(scheme, q, 0.17, O.DIFF_O4_EXPLICIT)
    np.testing.assert_allclose(mine, ref, rtol=0, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("n", [16, 40])
def test_o4_explicit_cores_closed_forms(n):
    """4th order: every row is exact on quartics (the 5-point interior and
    one-sided closures); periodic rows give the modified wavenumber
    (8 sin w - sin 2w) / 6 on sine waves."""
    h = 0.3
    x = np.arange(n) * h
    for p in range(5):
        f = (x - 1.1) ** p
        d = O.diff_core(O.OPEN, f[:, None], h, O.DIFF_O4_EXPLICIT)[:, 0]
        exact = p * (x - 1.1) ** (p - 1) if p else np.zeros(n)
        np.testing.assert_allclose(d, exact, atol=1e-9 * max(1.0, np.abs(exact).max()))
    # a quintic is not differentiated exactly: the closures are 4th order, not more
    f = (x - 1.1) ** 5
    d = O.diff_core(O.OPEN, f[:, None], h, O.DIFF_O4_EXPLICIT)[:, 0]
    assert np.abs(d - 5 * (x - 1.1) ** 4).max() > 1e-6
    hp = 2 * np.pi / n
    xp = np.arange(n) * hp
    for k in (1, 3, n // 4):
        f = np.sin(k * xp + 0.4)[:, None]
        d = O.diff_core(O.PERIODIC, f, hp, O.DIFF_O4_EXPLICIT)[:, 0]
        w = k * hp
        exact = (8 * np.sin(w) - np.sin(2 * w)) / 6 / hp * np.cos(k * xp + 0.4)
        np.testing.assert_allclose(d, exact, atol=1e-12 * k / hp)


@need_ref
@pytest.mark.parametrize("scheme", [O.PERIODIC, O.OPEN])
@pytest.mark.parametrize("n", [8, 24, 64])
def test_pade_cores_vs_reference_fortran(scheme, n):
    """diff_pen_core_o4_i_c_{p,o}_pade restated in the oracle against the reference's own."""
    q = _rand_lines(n, 5, 5 * n)
    ref = refproc.run("core", "diff_pade", scheme, q, 0.17)
    mine = O.diff_core(scheme, q, 0.17, O.DIFF_O4_PADE)
    np.testing.assert_allclose(mine, ref, rtol=0, atol=1e-13 * np.abs(ref).max())


@need_ref
@pytest.mark.parametrize("n", [8, 24, 64])
def test_o2_core_vs_reference_fortran(n):
    """diff_pen_core_o2_e_c_o restated in the oracle against the reference's own."""
    q = _rand_lines(n, 5, 7 * n)
    ref = refproc.run("core", "diff_o2", 1, q, 0.21)
    mine = O.diff_core(O.OPEN, q, 0.21, O.DIFF_O2_EXPLICIT)
    np.testing.assert_allclose(mine, ref, rtol=0, atol=1e-14 * np.abs(ref).max())


def test_o2_core_closed_form():
    """Exact on quadratics in every row (central and one-sided), not on cubics."""
    n, h = 20, 0.3
    x = np.arange(n) * h
    d = O.diff_core(O.OPEN, (1.5 - 0.7 * x + 0.4 * x ** 2)[:, None], h, O.DIFF_O2_EXPLICIT)[:, 0]
    np.testing.assert_allclose(d, -0.7 + 0.8 * x, atol=1e-12)
    d3 = O.diff_core(O.OPEN, (x ** 3)[:, None], h, O.DIFF_O2_EXPLICIT)[:, 0]
    assert np.abs(d3 - 3 * x ** 2).max() > 1e-3


@pytest.mark.parametrize("n", [16, 48])
def test_pade_periodic_closed_form(n):
    """Periodic Pade: the modified wavenumber 3 sin w / (2 + cos w) on sine
    waves, D(const) = 0."""
    h = 2 * np.pi / n
    x = np.arange(n) * h
    for k in (1, 3, n // 4):
        f = np.sin(k * x + 0.2)[:, None]
        d = O.diff_core(O.PERIODIC, f, h, O.DIFF_O4_PADE)[:, 0]
        w = k * h
        exact = 3 * np.sin(w) / (2 + np.cos(w)) / h * np.cos(k * x + 0.2)
        np.testing.assert_allclose(d, exact, atol=1e-12 * k / h)
    assert np.abs(O.diff_core(O.PERIODIC, np.ones((n, 2)), h, O.DIFF_O4_PADE)).max() < 1e-13


def _lele_modified_wavenumber(w):
    # Lele (1992) eq. (2.1.x): w' = (a sin w + b/2 sin 2w) / (1 + 2 alpha cos w),
    # alpha = 1/3, a = 14/9, b = 1/9
    return (14.0 / 9.0 * np.sin(w) + 1.0 / 18.0 * np.sin(2 * w)) / (1.0 + 2.0 / 3.0 * np.cos(w))


@pytest.mark.parametrize("n", [16, 48, 64])
def test_periodic_derivative_closed_form(n):
    h = 2 * np.pi / n
    x = np.arange(n) * h
    for k in (1, 3, n // 4):
        f = np.sin(k * x + 0.3)[:, None]
        d = O.diff_core(O.PERIODIC, f, h)[:, 0]
        exact = _lele_modified_wavenumber(k * h) / h * np.cos(k * x + 0.3)
        np.testing.assert_allclose(d, exact, atol=1e-12 * k)
    # D(const) = 0 and sum_i D f_i = 0
    f = _rand_lines(n, 3, 7)
    assert np.abs(O.diff_core(O.PERIODIC, np.ones((n, 2)), h)).max() < 1e-13
    assert np.abs(O.diff_core(O.PERIODIC, f, h).sum(axis=0)).max() < 1e-11


@pytest.mark.parametrize("n", [16, 40])
def test_open_derivative_polynomials(n):
    # The open scheme is exact for polynomials of degree <= 3 (4th-order closures).
    h = 0.37
    x = np.arange(n) * h
    for p in range(4):
        f = (x ** p)[:, None]
        d = O.diff_core(O.OPEN, f, h)[:, 0]
        exact = p * x ** (p - 1) if p > 0 else np.zeros(n)
        np.testing.assert_allclose(d, exact, atol=1e-10 * max(1.0, x.max() ** p))


@pytest.mark.parametrize("scheme", [O.PERIODIC, O.OPEN])
def test_filter_constant_and_transfer(scheme):
    n = 32
    c = O.filter_core(scheme, np.full((n, 2), 3.25), 0.35)
    np.testing.assert_allclose(c, 3.25, rtol=1e-14)
    if scheme == O.PERIODIC:
        # GV2000 transfer function T(w) = sum_k a_k cos(k w) / (1 + 2 alpha cos w)
        a = 0.35
        a0 = 11 / 16 + 5 * a / 8
        a1 = 15 / 32 + 17 * a / 16
        a2 = -3 / 16 + 3 * a / 8
        a3 = 1 / 32 - a / 16
        x = np.arange(n) * 2 * np.pi / n
        for k in (1, 5, 16):
            w = k * 2 * np.pi / n
            T = (a0 + a1 * np.cos(w) + a2 * np.cos(2 * w) + a3 * np.cos(3 * w)) / (1 + 2 * a * np.cos(w))
            f = np.cos(k * x)[:, None]
            np.testing.assert_allclose(O.filter_core(scheme, f, a)[:, 0], T * np.cos(k * x), atol=1e-13)


# --------------------------------------------------------------------------- closures / BC
@need_ref
@pytest.mark.parametrize("visc", ["sutherland"])
def test_closures_vs_reference(visc):
    rng = np.random.default_rng(3)
    n = 50
    q = np.asfortranarray(np.stack([1 + 0.2 * rng.random(n)] + [rng.standard_normal(n)] * 3 +
                                   [300 + 20 * rng.random(n)], axis=1)).reshape(n, 1, 1, 5, order="F")
    pars = dict(gamma=1.4, R=8.3144598, W=0.0289647, Pr=0.71, mu0=1.8e-5, T0=273.0)
    T, mu, lam = refproc.run("trafos", np.asfortranarray(q), *pars.values())
    cfg = O.Cfg((n, 1, 1), (1, 1, 1), **pars)
    T2, mu2, lam2 = O.point_closures(cfg, q[:, 0, 0, 0], q[:, 0, 0, 4])
    np.testing.assert_allclose(T2, T, rtol=1e-15)
    np.testing.assert_allclose(mu2, mu, rtol=1e-15)
    np.testing.assert_allclose(lam2, lam, rtol=1e-15)


def _boundary_states(n, seed):
    rng = np.random.default_rng(seed)
    q = np.zeros((n, 5), order="F")
    q[:, 0] = 1 + 0.3 * rng.random(n)
    q[:, 1:4] = rng.standard_normal((n, 3)) * q[:, :1]
    q[:, 4] = 0.5 + rng.random(n)
    return q


@need_ref
@pytest.mark.parametrize("k", [(0, -1, 0), (0, 1, 0), (1, 0, 0), (0.6, 0, -0.8)])
def test_characteristic_vs_reference(k):
    n, gamma = 64, 1.4
    q = _boundary_states(n, 1)
    aim = _boundary_states(n, 2)
    ref = refproc.run("characteristic", "q", q.reshape(n, 1, 1, 5, order="F"),
                      aim.reshape(n, 1, 1, 5, order="F"), k, gamma)
    mine = O.characteristic_q(q, aim, k, gamma)
    np.testing.assert_allclose(mine, ref.reshape(n, 5, order="F"), rtol=1e-14, atol=1e-14)
    qs = np.asfortranarray(np.random.default_rng(5).standard_normal((n, 5)))
    ref = refproc.run("characteristic", "qs", qs.reshape(n, 1, 1, 5, order="F"),
                      q.reshape(n, 1, 1, 5, order="F"), k, gamma)
    mine = O.characteristic_qs(qs, q, k, gamma)
    np.testing.assert_allclose(mine, ref.reshape(n, 5, order="F"), rtol=1e-14, atol=1e-14)


def test_characteristic_fixed_point():
    # aim = q leaves q unchanged; adjoint with qs = 0 stays 0
    q = _boundary_states(16, 9)
    np.testing.assert_allclose(O.characteristic_q(q, q, (0, 1, 0), 1.4), q, rtol=1e-14)
    z = np.zeros((16, 5), order="F")
    assert np.abs(O.characteristic_qs(z, q, (0, 1, 0), 1.4)).max() == 0


@need_ref
def test_sponge_vs_reference():
    n, gamma = 40, 1.4
    rng = np.random.default_rng(4)
    shape = np.asfortranarray(rng.random(n)).reshape(n, 1, 1, order="F")
    q = _boundary_states(n, 6).reshape(n, 1, 1, 5, order="F")
    qref = _boundary_states(n, 7).reshape(n, 1, 1, 5, order="F")
    y = np.asfortranarray(rng.standard_normal((n, 1, 1, 5)))
    cfg = O.Cfg((n, 1, 1), (1, 1, 1), bc=[(400, 1, n, 1, 1, 1, 1)])
    # sponge_rhs: rhs - shape*(q - qref)
    ref = refproc.run("sponge", 0, y, shape, q, qref, gamma)
    mine = O.bc_rhs(cfg, y, q, qref, shape)
    np.testing.assert_allclose(mine, ref, rtol=1e-15, atol=1e-15)
    # sponge_q (type 401): q - shape*(q - qref)
    cfg = O.Cfg((n, 1, 1), (1, 1, 1), bc=[(401, 1, n, 1, 1, 1, 1)])
    ref = refproc.run("sponge", 1, q, shape, q, qref, gamma)
    np.testing.assert_allclose(O.bc_q(cfg, q, qref, shape), ref, rtol=1e-15, atol=1e-15)
    # sponge_qs (type 401 on qs)
    ref = refproc.run("sponge", 2, y, shape, q, qref, gamma)
    np.testing.assert_allclose(O.bc_qs(cfg, y, q, shape), ref, rtol=1e-15, atol=1e-15)


# --------------------------------------------------------------------------- composite RHS
def _smooth_state(n, mach=0.3, gamma=1.4):
    x = np.arange(n) * 2 * np.pi / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    rho = 1 + 0.1 * np.sin(X + 0.2) * np.cos(Y) + 0.05 * np.cos(Z - 0.4)
    u = np.sin(X) * np.cos(Y) * np.cos(Z) + 0.2
    v = -np.cos(X) * np.sin(Y) * np.cos(Z) + 0.1 * np.sin(Z)
    w = 0.3 * np.sin(Y + Z) - 0.1
    p = 1 / (gamma * mach ** 2) + 0.2 * np.cos(X - Y) + 0.1 * np.sin(2 * Z)
    q = np.zeros((n, n, n, 5), order="F")
    q[..., 0], q[..., 1], q[..., 2], q[..., 3], q[..., 4] = rho, rho * u, rho * v, rho * w, p
    return q


def _spectral_rhs(q, cfg_args, n):
    """The same PDE evaluated with FFT derivatives (an independent discretisation).
    The heat flux carries the reference's dx^2 factor per direction (unit-spacing Dx calls)."""
    g = cfg_args["gamma"]
    h = 2 * np.pi / n
    k = np.fft.fftfreq(n, d=1.0 / n)

    def D(f, d):
        sh = [1, 1, 1]
        sh[d] = n
        return np.real(np.fft.ifft(1j * k.reshape(sh) * np.fft.fft(f, axis=d), axis=d))

    rho, p = q[..., 0], q[..., 4]
    u = [q[..., i + 1] / rho for i in range(3)]
    T = cfg_args["W"] / cfg_args["R"] * p / rho
    if cfg_args.get("visc_law", O.SUTHERLAND) == O.POWER:
        mu = cfg_args["mu0"] * (T / cfg_args["T0"]) ** cfg_args["visc_exp"]
    else:
        mu = cfg_args["mu0"] * (T / cfg_args["T0"]) ** 1.5 * ((cfg_args["T0"] + 110.4) / (T + 110.4))
    cp = g * (cfg_args["R"] / cfg_args["W"]) / (g - 1)
    lam = mu * cp / cfg_args["Pr"]
    G = [[D(u[i], j) for j in range(3)] for i in range(3)]  # G[i][j] = d u_i / d x_j
    div = G[0][0] + G[1][1] + G[2][2]
    tau = [[mu * (G[i][j] + G[j][i]) - (2.0 / 3.0) * mu * div * (i == j) for j in range(3)] for i in range(3)]
    out = np.zeros_like(q)
    out[..., 0] = -sum(D(rho * u[j], j) for j in range(3))
    for i in range(3):
        out[..., i + 1] = (-sum(D(rho * u[i] * u[j], j) for j in range(3)) - D(p, i)
                           + sum(D(tau[i][j], j) for j in range(3)))
    heat = sum(h * h * D(lam * D(T, j), j) for j in range(3))
    diss = sum(tau[i][j] * G[i][j] for i in range(3) for j in range(3))
    out[..., 4] = (-g * sum(D(p * u[j], j) for j in range(3)) + (g - 1) * sum(u[j] * D(p, j) for j in range(3))
                   + (g - 1) * (heat + diss))
    return out


@pytest.mark.parametrize("law", ["sutherland", "power"])
def test_rhs_direct_converges_to_spectral(law):
    """With the power law (the periodic-box workloads' viscosity) the
    spectral side evaluates mu0 (T/T0)**0.76 with numpy's power, so a wrong
    exponent, reference temperature or law would leave an O(1) error."""
    args = dict(gamma=1.4, R=1.0, W=1.0, Pr=0.71, mu0=0.05, T0=1.0)
    if law == "power":
        args.update(visc_law=O.POWER, visc_exp=0.76, T0=0.8)
    errs = []
    for n in (16, 32):
        h = 2 * np.pi / n
        cfg = O.Cfg((n, n, n), (h, h, h), **args)
        q = _smooth_state(n)
        fd = O.rhs_direct(cfg, q)
        sp = _spectral_rhs(q, args, n)
        errs.append([np.abs(fd[..., v] - sp[..., v]).max() / np.abs(sp[..., v]).max() for v in range(5)])
    errs = np.array(errs)
    # every component converges at (close to) 6th order; a dropped term or wrong
    # sign would leave an O(1) error that does not shrink
    assert np.all(errs[1] < 1e-5), errs
    assert np.all(errs[0] / errs[1] > 30), errs


def test_rhs_direct_invariants():
    n = 12
    h = 2 * np.pi / n
    cfg = O.Cfg((n, n, n), (h, h, h), mu0=0.01)
    # quiescent uniform state -> rhs = 0
    q = np.zeros((n, n, n, 5), order="F")
    q[..., 0] = 1.3
    q[..., 4] = 2.0
    assert np.abs(O.rhs_direct(cfg, q)).max() < 1e-12
    # uniform translation of a uniform state -> rhs = 0
    q[..., 1] = 1.3 * 0.4
    assert np.abs(O.rhs_direct(cfg, q)).max() < 1e-12
    # periodic mass and momentum conservation: sum of rhs over the box is 0
    q = _smooth_state(n)
    r = O.rhs_direct(cfg, q)
    for v in range(4):
        assert abs(r[..., v].sum()) < 1e-9 * np.abs(r[..., v]).sum()


def test_adjoint_duality():
    """The reference's adjoint is the continuous Euler adjoint of the system in
    U = (rho, rho u, rho v, rho w, p/(gamma-1)) (the (A^T)^-1 matrices quoted at
    src/equations/ns_rho_rhou_p.f90:310-372): d/dt <qs, dU> = 0, i.e.
    <rhs_adj(qs;q), D dq> + <qs, D L(q) dq> = 0 with D = diag(1,1,1,1,1/(gamma-1))
    and L the linearisation of the inviscid direct RHS (mu0 = 0).  With periodic
    compact derivatives (an antisymmetric operator) the identity holds to rounding."""
    n = 24
    h = 2 * np.pi / n
    gamma = 1.4
    cfg = O.Cfg((n, n, n), (h, h, h), mu0=0.0, gamma=gamma)
    q = _smooth_state(n, mach=0.5)
    x = np.arange(n) * h
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    qs = np.zeros_like(q)
    dq = np.zeros_like(q)
    for v in range(5):
        qs[..., v] = np.cos(X + v) * np.sin(Y - 0.5 * v) + 0.3 * np.cos(Z + v)
        dq[..., v] = np.sin(X - v) * np.cos(2 * Y) + 0.2 * np.sin(Z + 2 * v)
    eps = 1e-6
    Ldq = (O.rhs_direct(cfg, q + eps * dq) - O.rhs_direct(cfg, q - eps * dq)) / (2 * eps)
    ra = O.rhs_adjoint(cfg, qs, q)
    D = np.array([1, 1, 1, 1, 1 / (gamma - 1)])
    a = np.sum(ra * dq * D)
    b = np.sum(qs * Ldq * D)
    assert abs(a + b) < 1e-9 * abs(a), (a, b)
    # the unweighted pairing does not hold: the weighting is part of the definition
    assert abs(np.sum(ra * dq) + np.sum(qs * Ldq)) > 0.1 * abs(a)


def test_rk4_amplification_with_sponge():
    """Uniform state relaxed by a sponge of strength s: the RHS is -s (q - qref),
    so one RK4 step multiplies q - qref by 1 - z + z^2/2 - z^3/6 + z^4/24, z = s dt."""
    n = 6
    s, dt = 3.0, 0.07
    cfg = O.Cfg((n, n, n), (1, 1, 1), dt=dt, mu0=0.0, bc=[(400, 1, n, 1, n, 1, n)])
    q = np.zeros((n, n, n, 5), order="F")
    q[..., 0], q[..., 4] = 1.2, 3.0
    qref = q.copy(order="F")
    qref[..., 0], qref[..., 4] = 1.0, 2.5
    shape = np.full((n, n, n), s, order="F")
    q1 = O.time_stepper_direct(cfg, q, qref=qref, shape=shape)
    z = s * dt
    amp = 1 - z + z ** 2 / 2 - z ** 3 / 6 + z ** 4 / 24
    np.testing.assert_allclose(q1 - qref, amp * (q - qref), rtol=1e-13, atol=1e-14)
