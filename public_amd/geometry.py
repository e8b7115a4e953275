"""Host-side geometry of a run: the grids, the metric of a deformed grid and
the sponge shape, in the form the reference's initialisation leaves them.

* `cartesian_grid` restates cartesian_2_cartesian_strech_shift
  (src/geometry/grid_transform.f90:20-121): the calculation space, with
  dx = L/n on periodic and L/(n-1) on open directions.
* `tanh_stretch_y` is a case-level distortion (the reference's
  distort_user_defined hook, src/geometry/cartesian_single_block.f90:167-176):
  y clustered around its mid-plane, s = tanh(beta y') / tanh(beta) uniform,
  with clustering factor beta (mid-plane spacing / end spacing = 1/cosh^2 beta).
* `calc_metric` follows grid_metric.f90:16-188 (Visbal & Gaitonde's
  conservative metric with the periodic corrections) and takes every compact
  derivative on the GPU through the library (`Cruna.Dx`).
* `sponge_shape` follows shape_functions.f90:69-150 (quadratic or cosine
  ramps; a negative length disables a direction).
"""
import numpy as np

from .cruna import PERIODIC


def cartesian_grid(n, x0, x1, scheme):
    """(X, dx): X (n1,n2,n3,3) Fortran order, dx per direction."""
    vecs, dx = [], []
    for d in range(3):
        h = (x1[d] - x0[d]) / (n[d] - (0 if scheme[d] == PERIODIC else 1))
        dx.append(h)
        vecs.append(x0[d] + np.arange(n[d]) * h if n[d] > 1 else np.full(1, float(x0[d])))
    X = np.stack(np.meshgrid(*vecs, indexing="ij"), axis=-1)
    return np.asfortranarray(X), tuple(dx)


def tanh_stretch_y(X_calc, y0, y1, beta):
    """Physical grid: y = ym + (L/2) atanh(s tanh(beta)) / beta for the uniform
    s = (y_calc - ym) / (L/2) in [-1, 1] (fine at the mid-plane, coarse at the ends)."""
    X = np.array(X_calc, order="F", copy=True)
    ym, half = 0.5 * (y0 + y1), 0.5 * (y1 - y0)
    s = np.clip((X_calc[..., 1] - ym) / half, -1.0, 1.0)
    X[..., 1] = ym + half * np.arctanh(s * np.tanh(beta)) / beta
    return X


def calc_metric(ctx, X, X_calc):
    """X_metric (n1,n2,n3,3,3) and X_jacobian (n1,n2,n3) of grid_metric::calc_metric.

    ctx: a `Cruna` context of this grid (its schemes set the periodic
    corrections c1..c3, its dx the derivative spacing)."""
    cfg = ctx.cfg
    c = [1.0 if cfg.scheme[d] == PERIODIC else 0.0 for d in range(3)]
    dx = [cfg.dx[d] for d in range(3)]
    Xt = [np.asfortranarray(X[..., d] - c[d] * X_calc[..., d]) for d in range(3)]

    def D(d, f):
        return ctx.Dx(d, np.asfortranarray(f), dx[d])

    # d[a][b] = d x_{a+1} / d xi_{b+1}
    dd = [[D(b, Xt[a]) for b in range(3)] for a in range(3)]
    x1, x2, x3 = Xt
    d11, d12, d13 = dd[0]
    d21, d22, d23 = dd[1]
    d31, d32, d33 = dd[2]
    c1, c2, c3 = c
    M = np.zeros(X.shape[:3] + (3, 3), order="F")
    M[..., 0, 0] = D(2, d22 * x3) - D(1, d23 * x3) + c2 * d33 + c3 * d22 + c2 * c3
    M[..., 0, 1] = D(1, d13 * x3) - D(2, d12 * x3) - c3 * d12
    M[..., 0, 2] = D(2, d12 * x2) - D(1, d13 * x2) - c2 * d13
    M[..., 1, 0] = D(0, d23 * x3) - D(2, d21 * x3) - c3 * d21
    M[..., 1, 1] = D(2, d11 * x3) - D(0, d13 * x3) + c1 * d33 + c3 * d11 + c1 * c3
    M[..., 1, 2] = D(0, d13 * x2) - D(2, d11 * x2) - c1 * d23
    M[..., 2, 0] = D(1, x3 * d21) - D(0, x3 * d22) - c2 * d31
    M[..., 2, 1] = D(0, d12 * x3) - D(1, d11 * x3) - c1 * d32
    M[..., 2, 2] = D(1, d11 * x2) - D(0, d12 * x2) + c1 * d22 + d11 * c2 + c1 * c2
    J = M[..., 2, 0] * d13 + M[..., 2, 1] * d23 + M[..., 2, 2] * (d33 + c3)
    return M, np.asfortranarray(1.0 / J)


def sponge_shape(X, x0, x1, length, amplitude, cos_shape=False):
    """sigma (n1,n2,n3) of init_general_shape_fct for the box [x0, x1] and
    sponge lengths per direction (negative: no sponge in that direction)."""
    phi = np.ones(X.shape[:3])

    def ramp(x, a, ln):
        if cos_shape:
            return 0.5 + 0.5 * np.cos((x - a) * (np.pi / ln))
        return 1.0 - ((x - a) / ln) ** 2

    for d in range(3):
        x = X[..., d]
        lo = x0[d] + length[d]
        s = ramp(x, lo, length[d])
        s[x > lo] = 1.0
        phi = phi * s
        hi = x1[d] - length[d]
        s = ramp(x, hi, length[d])
        s[x < hi] = 1.0
        phi = phi * s
    return np.asfortranarray(amplitude * (1.0 - phi))


def cfl_dt(q, X, cfl=0.5, gamma=1.4):
    """dt = cfl * min spacing / max(|u| + c) (check_state.f90:125-136 with the
    smallest physical spacing of the grid)."""
    rho = q[..., 0]
    u = np.sqrt(q[..., 1] ** 2 + q[..., 2] ** 2 + q[..., 3] ** 2) / rho
    c = np.sqrt(gamma * q[..., 4] / rho)
    hmin = np.inf
    for d in range(3):
        if X.shape[d] > 1:
            diff = np.diff(X[..., d], axis=d)
            hmin = min(hmin, float(np.abs(diff).min()))
    return cfl * hmin / float((u + c).max())
