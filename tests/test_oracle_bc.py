"""Pins of the oracle's RHS boundary rows: characteristic_rhs_k1d (type 204,
src/boundaries/nonreflecting.f90:237-341) and the slip wall (type 100).

The wave transform and its back-transform must be inverses: when every wave
enters the domain nothing is removed, and the new RHS is the closed form
rho_t = (rho_aim - rho)/dt + rho_t, u_t = ((u_aim - u)/dt + m_t/rho) (rho + d rho)
(momentum slots), p_t = (p_aim - p)/dt + p_t, for any unit normal; when
every wave leaves it, the RHS is zero; aim = q with a zero RHS stays zero."""
import numpy as np
import pytest

from oracle import oracle as O

GAMMA, DT = 1.4, 2e-3


def _states(m, seed, speed):
    rng = np.random.default_rng(seed)
    q = np.zeros((m, 5), order="F")
    rho = 1 + 0.1 * rng.random(m)
    p = 1 / 1.4 * (1 + 0.1 * rng.random(m))
    c = np.sqrt(GAMMA * p / rho)
    return q, rho, p, c, rng


def _conservative(rho, u, p):
    q = np.zeros((rho.size, 5), order="F")
    q[:, 0] = rho
    q[:, 1:4] = rho[:, None] * u
    q[:, 4] = p
    return q


@pytest.mark.parametrize("k", [(0.0, -1.0, 0.0), (0.6, 0.0, 0.8), (-0.48, 0.6, 0.64)])
def test_all_incoming_keeps_every_wave(k):
    m = 16
    _, rho, p, c, rng = _states(m, 1, 0)
    k = np.asarray(k) / np.linalg.norm(k)
    # velocity along -k (into the domain) faster than sound: every eigenvalue of (-k).u >= 0
    u = (-k)[None, :] * (c * 1.5)[:, None] + 0.05 * rng.standard_normal((m, 3))
    q = _conservative(rho, u, p)
    aim = _conservative(rho * 1.02, u * 0.97, p * 1.01)
    rhs = np.asfortranarray(rng.standard_normal((m, 5)) * 0.1)
    got = O.characteristic_rhs(rhs, q, aim, k, GAMMA, DT)
    d_rho = aim[:, 0] - rho + rhs[:, 0] * DT
    exp = np.zeros_like(got)
    exp[:, 0] = d_rho / DT
    u_aim = aim[:, 1:4] / aim[:, :1]
    du = (u_aim - u) + rhs[:, 1:4] * (DT / rho)[:, None]
    exp[:, 1:4] = du * ((rho + d_rho) / DT)[:, None]
    exp[:, 4] = (aim[:, 4] - p + rhs[:, 4] * DT) / DT
    np.testing.assert_allclose(got, exp, rtol=1e-11, atol=1e-11)


def test_all_outgoing_removes_every_wave():
    m = 8
    _, rho, p, c, rng = _states(m, 2, 0)
    k = np.array([1.0, 0.0, 0.0])
    u = k[None, :] * (c * 1.3)[:, None]  # along +k: out of the domain, supersonic
    q = _conservative(rho, u, p)
    aim = _conservative(rho * 1.05, u * 0.9, p * 1.02)
    rhs = np.asfortranarray(rng.standard_normal((m, 5)))
    got = O.characteristic_rhs(rhs, q, aim, k, GAMMA, DT)
    assert np.abs(got).max() < 1e-12


def test_aim_equal_to_state_and_zero_rhs_stays_zero():
    m = 8
    _, rho, p, c, rng = _states(m, 3, 0)
    u = 0.3 * rng.standard_normal((m, 3))
    q = _conservative(rho, u, p)
    got = O.characteristic_rhs(np.zeros((m, 5), order="F"), q, q, (0.0, 0.0, -1.0), GAMMA, DT)
    assert np.abs(got).max() < 1e-13


def test_subsonic_outflow_keeps_only_the_upstream_acoustic_wave():
    """Subsonic outflow through +x: the inward normal is -x, so lamS = -u < 0
    and lamM = -u - c < 0 remove R1..R3 and R5, and only R4 (lamP = c - u > 0)
    survives: the new RHS is R4's eigenvector (rho/(2c), -1/2, 0, 0, rho c/2)
    times R4 over dt (momentum: times rho + d rho)."""
    m = 6
    _, rho, p, c, rng = _states(m, 4, 0)
    u = np.zeros((m, 3))
    u[:, 0] = 0.4 * c
    q = _conservative(rho, u, p)
    aim = _conservative(rho * 1.01, u * 1.03, p * 0.99)
    rhs = np.asfortranarray(rng.standard_normal((m, 5)) * 0.05)
    got = O.characteristic_rhs(rhs, q, aim, (1.0, 0.0, 0.0), GAMMA, DT)
    du = aim[:, 1:4] / aim[:, :1] - u + rhs[:, 1:4] * (DT / rho)[:, None]
    dp = aim[:, 4] - p + rhs[:, 4] * DT
    R4 = -du[:, 0] + dp / (rho * c)
    d_rho = rho / (2 * c) * R4
    exp = np.zeros_like(got)
    exp[:, 0] = d_rho / DT
    exp[:, 1] = -0.5 * R4 * (rho + d_rho) / DT
    exp[:, 4] = 0.5 * rho * c * R4 / DT
    np.testing.assert_allclose(got, exp, rtol=1e-11, atol=1e-12)


def test_wall_and_characteristic_rows_in_the_rhs():
    """Rows 100 and 204 act on the finished RHS through orc_bc_rhs, in row order,
    only inside their boxes."""
    n = (16, 16, 16)
    rng = np.random.default_rng(5)
    q = np.zeros(n + (5,), order="F")
    q[..., 0] = 1 + 0.05 * rng.random(n)
    q[..., 1:4] = 0.1 * rng.standard_normal(n + (3,))
    q[..., 4] = 1 / 1.4
    rhs = np.asfortranarray(rng.standard_normal(n + (5,)))
    cfg = O.Cfg(n, (0.1,) * 3, dt=DT, bc=[(100, 1, 16, 1, 1, 1, 16, 3), (204, 1, 16, 16, 16, 1, 16, 0, 1, 0)])
    got = O.bc_rhs(cfg, rhs, q, qref=q)
    assert np.array_equal(got[:, 0, :, 2], np.zeros((16, 16)))
    assert np.array_equal(got[:, 1:15], rhs[:, 1:15])
    face = O.characteristic_rhs(rhs[:, 15].reshape(-1, 5, order="F"), q[:, 15].reshape(-1, 5, order="F"),
                                q[:, 15].reshape(-1, 5, order="F"), (0, 1, 0), 1.4, DT)
    np.testing.assert_allclose(got[:, 15].reshape(-1, 5, order="F"), face, rtol=0, atol=0)


@pytest.mark.parametrize("face", ["lo", "hi"])
def test_zero_gradient_row_is_the_characteristic_row_toward_the_inner_neighbour(face):
    """Type 200 (characteristic_zero_grad_init + characteristic_q/_qs) equals
    type 201 with the target (q) or the linearisation state (qs) taken one
    point inside, against the normal: the index shift of the zero-gradient
    aim, pinned on both y faces."""
    n = (16, 16, 16)
    rng = np.random.default_rng(7)
    q = np.zeros(n + (5,), order="F")
    q[..., 0] = 1 + 0.05 * rng.random(n)
    q[..., 1:4] = 0.2 * rng.standard_normal(n + (3,))
    q[..., 4] = 1 / 1.4 * (1 + 0.05 * rng.random(n))
    j, k2, jn = (1, -1.0, 2) if face == "lo" else (16, 1.0, 15)
    row = (1, 16, j, j, 1, 16, 0, k2, 0)
    shifted = q.copy(order="F")
    shifted[:, j - 1] = q[:, jn - 1]
    got = O.bc_q(O.Cfg(n, (0.1,) * 3, bc=[(200,) + row]), q)
    ref = O.bc_q(O.Cfg(n, (0.1,) * 3, bc=[(201,) + row]), q, qref=shifted)
    assert np.array_equal(got, ref)
    assert np.abs(got - q).max() > 1e-3
    qs = np.asfortranarray(rng.standard_normal(n + (5,)))
    got_s = O.bc_qs(O.Cfg(n, (0.1,) * 3, bc=[(200,) + row]), qs, q)
    ref_s = O.bc_qs(O.Cfg(n, (0.1,) * 3, bc=[(201,) + row]), qs, shifted)
    assert np.array_equal(got_s, ref_s)


def test_per_point_normal_rows_match_their_row_normal_forms():
    """Types 202 / 205 (characteristic_*_k4d with kvector(box,:)) equal 201 /
    204 applied point by point with each point's normal as the row normal;
    with a constant field equal to the row's normal they equal 201 / 204."""
    n = (16, 16, 16)
    rng = np.random.default_rng(9)
    q = np.zeros(n + (5,), order="F")
    q[..., 0] = 1 + 0.05 * rng.random(n)
    q[..., 1:4] = 0.2 * rng.standard_normal(n + (3,))
    q[..., 4] = 1 / 1.4 * (1 + 0.05 * rng.random(n))
    qref = q.copy(order="F")
    qref[..., 0] *= 1.02
    kv = np.zeros(n + (3,), order="F")
    kv[..., 1] = -1.0
    kv[..., 0] = 0.3 * np.sin(np.arange(16))[:, None, None]
    kv /= np.linalg.norm(kv, axis=-1, keepdims=True)
    box = (1, 16, 1, 1, 1, 16)
    got = O.bc_q(O.Cfg(n, (0.1,) * 3, bc=[(202,) + box], kvector=kv), q, qref=qref)
    ref = q.copy(order="F")
    for i in range(16):
        for kk in range(16):
            row = (201, i + 1, i + 1, 1, 1, kk + 1, kk + 1) + tuple(kv[i, 0, kk])
            ref = O.bc_q(O.Cfg(n, (0.1,) * 3, bc=[row]), ref, qref=qref)
    assert np.array_equal(got, ref)
    const = np.zeros(n + (3,), order="F")
    const[..., 1] = -1.0
    a = O.bc_q(O.Cfg(n, (0.1,) * 3, bc=[(202,) + box], kvector=const), q, qref=qref)
    b = O.bc_q(O.Cfg(n, (0.1,) * 3, bc=[(201,) + box + (0, -1, 0)]), q, qref=qref)
    assert np.array_equal(a, b)
    rhs = np.asfortranarray(rng.standard_normal(n + (5,)))
    a = O.bc_rhs(O.Cfg(n, (0.1,) * 3, dt=DT, bc=[(205,) + box], kvector=const), rhs, q, qref=qref)
    b = O.bc_rhs(O.Cfg(n, (0.1,) * 3, dt=DT, bc=[(204,) + box + (0, -1, 0)]), rhs, q, qref=qref)
    assert np.array_equal(a, b)
    qs = np.asfortranarray(rng.standard_normal(n + (5,)))
    a = O.bc_qs(O.Cfg(n, (0.1,) * 3, bc=[(202,) + box], kvector=const), qs, q)
    b = O.bc_qs(O.Cfg(n, (0.1,) * 3, bc=[(201,) + box + (0, -1, 0)]), qs, q)
    assert np.array_equal(a, b)


def _zonal_case(n, seed):
    rng = np.random.default_rng(seed)
    q = np.zeros(n + (5,), order="F")
    q[..., 0] = 1 + 0.05 * rng.random(n)
    q[..., 1:4] = 0.3 * rng.standard_normal(n + (3,))
    q[..., 4] = 1 / 1.4 * (1 + 0.05 * rng.random(n))
    qref = q.copy(order="F")
    qref[..., 0] *= 1.02
    qref[..., 4] *= 0.99
    qref[..., 1:4] += 0.05
    kv = np.asfortranarray(rng.standard_normal(n + (3,)))
    kv /= np.linalg.norm(kv, axis=-1, keepdims=True)
    shape = np.asfortranarray(rng.random(n))
    return q, qref, kv, shape, rng


def _prim(q):
    return np.concatenate([q[..., :1], q[..., 1:4] / q[..., :1], q[..., 4:]], axis=-1)


def test_zonal_rows_are_the_per_point_rows_over_the_domain_scaled_by_the_shape():
    """Types 203 / 206 (characteristic_*_k4d(q, qref, kvector, shape) on the
    whole domain, boundary_conditions.f90:93-108, 276-279): with shape 1 they
    equal 202 / 205 on a box covering the domain, whatever box columns the
    row holds; the change of the primitive state that 203 makes is linear in
    the shape (the wave mask depends on the state only), shape 0 leaves q,
    qs and the RHS as they are; 206 is affine in the shape between the RHS
    (shape 0) and 205's (shape 1); the adjoint 203 scales qs's change."""
    n = (12, 10, 8)
    q, qref, kv, shape, rng = _zonal_case(n, 21)
    dx = (0.1,) * 3
    whole = (1, n[0], 1, n[1], 1, n[2])
    one, zero = np.ones(n, order="F"), np.zeros(n, order="F")
    z = O.Cfg(n, dx, dt=DT, bc=[(203, 0, 0, 0, 0, 0, 0)], kvector=kv)
    p = O.Cfg(n, dx, dt=DT, bc=[(202,) + whole], kvector=kv)
    assert np.array_equal(O.bc_q(z, q, qref=qref, shape=one), O.bc_q(p, q, qref=qref))
    base = _prim(q)
    d1 = _prim(O.bc_q(z, q, qref=qref, shape=one)) - base
    ds = _prim(O.bc_q(z, q, qref=qref, shape=shape)) - base
    assert np.abs(ds - shape[..., None] * d1).max() < 1e-14 * np.abs(base).max()
    assert np.abs(d1).max() > 1e-3  # the rows do change the state
    assert np.allclose(O.bc_q(z, q, qref=qref, shape=zero), q, rtol=1e-15, atol=0)

    rhs = np.asfortranarray(rng.standard_normal(n + (5,)))
    z6 = O.Cfg(n, dx, dt=DT, bc=[(206, 0, 0, 0, 0, 0, 0)], kvector=kv)
    p5 = O.Cfg(n, dx, dt=DT, bc=[(205,) + whole], kvector=kv)
    r1 = O.bc_rhs(z6, rhs, q, qref=qref, shape=one)
    assert np.array_equal(r1, O.bc_rhs(p5, rhs, q, qref=qref))
    assert np.array_equal(O.bc_rhs(z6, rhs, q, qref=qref, shape=zero), rhs)
    rs = O.bc_rhs(z6, rhs, q, qref=qref, shape=shape)
    s = shape[..., None]
    assert np.abs(rs - ((1 - s) * rhs + s * r1)).max() < 1e-13 * np.abs(r1).max()

    qs = np.asfortranarray(rng.standard_normal(n + (5,)))
    a1 = O.bc_qs(z, qs, q, shape=one)
    assert np.array_equal(a1, O.bc_qs(p, qs, q))
    a_s = O.bc_qs(z, qs, q, shape=shape)
    assert np.abs((a_s - qs) - s * (a1 - qs)).max() < 1e-14 * np.abs(qs).max()


@pytest.mark.parametrize("flag", [-1, 0, 1])
def test_zonal_row_flag_picks_the_calls_it_acts_in(flag):
    """Column 8 of a 203 row (init_boundary_conditions,
    boundary_conditions.f90:434-438): -1 acts only in the call after the RK
    step (outside_rhs), 1 only in the RHS's own calls, 0 in both.  A step
    with flag 1 is the RK step of a flag-0 row; a step with flag -1 is the
    RK step without the row followed by the row once."""
    n = (12, 10, 8)
    q, qref, kv, shape, _ = _zonal_case(n, 22)
    q[..., 1:4] *= 0.3
    kw = dict(dt=DT, mu0=1e-2, filter_freq=(0, 0, 0), kvector=kv)
    dx = (0.1,) * 3
    row = (203, 0, 0, 0, 0, 0, 0, flag)
    got = O.run_direct(O.Cfg(n, dx, bc=[row], **kw), q, 1, qref=qref, shape=shape)
    both = O.Cfg(n, dx, bc=[(203, 0, 0, 0, 0, 0, 0, 0)], **kw)
    if flag == 1:
        ref = O.time_stepper_direct(both, q, qref=qref, shape=shape)
    else:
        inner = O.time_stepper_direct(both if flag == 0 else O.Cfg(n, dx, **kw), q, qref=qref, shape=shape)
        ref = O.bc_q(both, inner, qref=qref, shape=shape)
    assert np.array_equal(got, ref)
    other = O.run_direct(O.Cfg(n, dx, bc=[(203, 0, 0, 0, 0, 0, 0, -flag or 1)], **kw), q, 1,
                         qref=qref, shape=shape)
    assert not np.array_equal(got, other)


def test_adjoint_stops_on_characteristic_rhs_rows():
    """set_boundary_condition_rhss has its cases 204-206 commented out
    (boundary_conditions.f90:339-353): the adjoint stops on such rows."""
    n = (8, 8, 8)
    q, qref, kv, shape, rng = _zonal_case(n, 23)
    for row in [(204, 1, 8, 1, 1, 1, 8, 0, -1, 0), (205, 1, 8, 1, 1, 1, 8), (206,)]:
        cfg = O.Cfg(n, (0.1,) * 3, bc=[row], kvector=kv)
        with pytest.raises(ValueError, match="stops"):
            O.rhs_adjoint(cfg, q, q, shape=shape)


def test_adjoint_wall_row_zeroes_its_variable_after_everything_else():
    """Type 100 in set_boundary_condition_rhss (boundary_conditions.f90:319-320):
    the adjoint RHS with a wall row is the one without it, with the named
    variable zero on the row's box only, g included (the row acts last)."""
    n = (12, 10, 8)
    q, qref, kv, shape, rng = _zonal_case(n, 24)
    qs = np.asfortranarray(rng.standard_normal(n + (5,)))
    g = np.asfortranarray(rng.standard_normal(n + (5,)))
    kw = dict(dt=DT, scheme=(O.PERIODIC, O.OPEN, O.PERIODIC))
    base = O.rhs_adjoint(O.Cfg(n, (0.1,) * 3, **kw), qs, q, g=g)
    got = O.rhs_adjoint(O.Cfg(n, (0.1,) * 3, bc=[(100, 1, 12, 1, 1, 1, 8, 3)], **kw), qs, q, g=g)
    ref = base.copy(order="F")
    ref[:, 0, :, 2] = 0.0
    assert np.array_equal(got, ref)
    assert np.abs(base[:, 0, :, 2]).max() > 0
