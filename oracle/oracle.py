"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE.

Loads oracle/liboracle.so (the C restatement of the reference algorithm) and,
when built, oracle/_ref/libcruna_ref.so (the reference's own Fortran leaf
modules).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg import this module; the product library never does.

Arrays are numpy float64 in Fortran order, shape (n1, n2, n3[, 5]).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MAX_BC = 32

PERIODIC, OPEN = 0, 1
DIFF_O6_COMPACT, DIFF_O4_EXPLICIT, DIFF_O4_PADE, DIFF_O2_EXPLICIT = 0, 1, 2, 3
EQ_NS, EQ_EULER = 0, 1
SUTHERLAND, POWER = 0, 1


class OrcCfg(C.Structure):
    _fields_ = [
        ("n", C.c_int * 3),
        ("dx", C.c_double * 3),
        ("scheme", C.c_int * 3),
        ("filter_alpha", C.c_double),
        ("gamma", C.c_double),
        ("R", C.c_double),
        ("W", C.c_double),
        ("Pr", C.c_double),
        ("mu0", C.c_double),
        ("T0", C.c_double),
        ("visc_law", C.c_int),
        ("visc_exp", C.c_double),
        ("dt", C.c_double),
        ("filter_freq", C.c_int * 3),
        ("afilter_freq", C.c_int * 3),
        ("nbc", C.c_int),
        ("bc", (C.c_double * 10) * MAX_BC),
        ("X_metric", C.c_void_p),
        ("X_jacobian", C.c_void_p),
        ("diff_kind", C.c_int * 3),
        ("penal_phi", C.c_void_p),
        ("penal_inv_phi", C.c_void_p),
        ("wfilter_freq", C.c_int * 3),
        ("equations", C.c_int),
        ("kvector", C.c_void_p),
    ]


_lib = None
_ref = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="F_CONTIGUOUS")
_vp = C.c_void_p


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        P = C.POINTER(OrcCfg)
        L.orc_dgttrf.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _ip]
        L.orc_dgttrf.restype = C.c_int
        L.orc_dgttrs.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _ip, _dp, C.c_int]
        L.orc_diff_core.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_double]
        L.orc_diff_core_o4.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_double]
        L.orc_diff_core_pade.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_double]
        L.orc_diff_core_o2.argtypes = [C.c_int, C.c_int, _dp, C.c_double]
        L.orc_filter_core.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_double]
        L.orc_diff3d.argtypes = [P, C.c_int, _dp, _dp, C.c_double]
        L.orc_filter4d.argtypes = [P, C.c_int, _dp, C.c_int]
        for f in ("orc_T",):
            getattr(L, f).argtypes = [P, C.c_double, C.c_double]
            getattr(L, f).restype = C.c_double
        L.orc_mu.argtypes = [P, C.c_double]
        L.orc_mu.restype = C.c_double
        L.orc_lambda.argtypes = [P, C.c_double]
        L.orc_lambda.restype = C.c_double
        L.orc_rhs_direct.argtypes = [P, _dp, _dp, _vp, _vp, C.c_int]
        L.orc_calc_metric.argtypes = [P, _dp, _dp, _dp, _dp]
        L.orc_rhs_direct_gt.argtypes = [P, _dp, _dp, _dp, _dp, _vp, _vp]
        L.orc_rhs_direct_euler_gt.argtypes = [P, _dp, _dp, _dp, _dp, _vp, _vp]
        L.orc_rhs_direct_ev.argtypes = [P, _dp, _dp, _dp, _dp, _vp, _vp]
        L.orc_filter_weighted4d.argtypes = [P, C.c_int, _dp, C.c_int, _vp, _vp]
        L.orc_characteristic_rhs_k1d.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_double,
                                                 C.c_double]
        L.orc_rhs_adjoint.argtypes = [P, _dp, _dp, _dp, _vp, _vp]
        L.orc_bc_q.argtypes = [P, _dp, _vp, _vp]
        L.orc_bc_qs.argtypes = [P, _dp, _dp, _vp]
        L.orc_bc_rhs.argtypes = [P, _dp, _dp, _vp, _vp]
        L.orc_characteristic_q_k1d.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, C.c_double]
        L.orc_characteristic_qs_k1d.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, C.c_double]
        L.orc_sponge_shape.argtypes = [P, _dp, _dp, _dp, C.c_double, C.c_int, _dp, _dp, _dp, _dp]
        L.orc_time_stepper_direct.argtypes = [P, _dp, _dp, _vp, _vp, C.c_int]
        L.orc_time_stepper_adjoint.argtypes = [P, _dp, _dp, _dp, _vp, _vp]
        L.orc_run_direct.argtypes = [P, _dp, C.c_int, C.c_int, _vp, _vp, C.c_int]
        L.orc_run_adjoint.argtypes = [P, _dp, _dp, C.c_int, _vp]
        L.orc_set_core_hooks.argtypes = [_vp, _vp]
        L.orc_set_core_hooks.restype = None
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads.restype = None
        L.orc_run_direct_conv.argtypes = [P, _dp, C.c_int, C.c_int, C.c_int, C.c_double, _vp, _vp, C.c_int,
                                          C.POINTER(C.c_double), C.POINTER(C.c_int)]
        _lib = L
    return _lib


def ref_lib():
    """The reference's compiled leaf modules, or None when not built."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libcruna_ref.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        L.ref_init.argtypes = [C.c_double] * 7 + [C.c_int]
        L.ref_diff.argtypes = [C.c_int, _dp, C.c_int, C.c_int, C.c_double]
        L.ref_diff_o4.argtypes = [C.c_int, _dp, C.c_int, C.c_int, C.c_double]
        L.ref_diff_pade.argtypes = [C.c_int, _dp, C.c_int, C.c_int, C.c_double]
        L.ref_diff_o2.argtypes = [_dp, C.c_int, C.c_int, C.c_double]
        L.ref_filter.argtypes = [C.c_int, _dp, C.c_int, C.c_int]
        L.ref_trafos.argtypes = [_dp, C.c_int, _dp, _dp, _dp]
        L.ref_characteristic_q.argtypes = [_dp, _dp, C.c_int, _dp]
        L.ref_characteristic_qs.argtypes = [_dp, _dp, C.c_int, _dp]
        L.ref_sponge.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_int]
        L.ref_diff_omp.argtypes = [C.c_int, _dp, C.c_int, C.c_int, C.c_double]
        L.ref_set_threads.argtypes = [C.c_int]
        _ref = L
    return _ref


def use_reference_cores(threads=0, alpha=0.35):
    """Run the reference's own compiled o6 cores (oracle/_ref) inside the
    oracle's derivatives and filters: the serial ones, or with threads > 0
    the reference's OpenMP derivative core with that many threads (periodic
    lines; the filter has no threaded core in the reference).  Returns False
    when oracle/_ref is not built.  use_restated_cores() undoes it."""
    R = ref_lib()
    if R is None:
        return False
    R.ref_init(alpha, 1.4, 1.0, 1.0, 0.71, 0.0, 1.0, 1)
    if threads > 0:
        R.ref_set_threads(int(threads))
    diff = C.cast(R.ref_diff_omp if threads > 0 else R.ref_diff, C.c_void_p)
    lib().orc_set_core_hooks(diff, C.cast(R.ref_filter, C.c_void_p))
    return True


def use_restated_cores():
    lib().orc_set_core_hooks(None, None)


def set_threads(n):
    lib().orc_set_threads(int(n))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Cfg:
    """Python view of orc_cfg with the reference's parameter names."""

    def __init__(self, n, dx, scheme=(PERIODIC,) * 3, filter_alpha=0.35, gamma=1.4, R=1.0,
                 W=1.0, Pr=0.71, mu0=1.0 / 1600.0, T0=1.0, visc_law=SUTHERLAND, visc_exp=0.76,
                 dt=1e-3, filter_freq=(1, 1, 1), afilter_freq=(0, 0, 0), bc=(), metric=None,
                 diff_kind=(0, 0, 0), penalization=None, wfilter_freq=(0, 0, 0), equations=0,
                 kvector=None):
        """metric: (X_metric (n1,n2,n3,3,3), X_jacobian (n1,n2,n3)) of a deformed
        grid; the direct RHS is then the grid-transform one.  diff_kind per
        direction: DIFF_O6_COMPACT, DIFF_O4_EXPLICIT or DIFF_O4_PADE.
        penalization: (phi, 1/phi) (n1,n2,n3) each; the direct RHS is then
        ns_rho_rhou_p_ev_phi's."""
        c = OrcCfg()
        for d in range(3):
            c.n[d] = int(n[d])
            c.dx[d] = float(dx[d])
            c.scheme[d] = int(scheme[d])
            c.filter_freq[d] = int(filter_freq[d])
            c.afilter_freq[d] = int(afilter_freq[d])
            c.diff_kind[d] = int(diff_kind[d])
            c.wfilter_freq[d] = int(wfilter_freq[d])
        c.filter_alpha = filter_alpha
        c.gamma, c.R, c.W, c.Pr, c.mu0, c.T0 = gamma, R, W, Pr, mu0, T0
        c.visc_law, c.visc_exp, c.dt = visc_law, visc_exp, dt
        c.equations = int(equations)
        if c.equations == EQ_EULER and penalization is not None and metric is None:
            raise ValueError("the reference has no penalised Euler module for Cartesian grids")
        c.nbc = len(bc)
        for r, row in enumerate(bc):
            for j in range(10):
                c.bc[r][j] = float(row[j]) if j < len(row) else 0.0
        self._metric = None
        if metric is not None:
            Xm, Xj = _f(metric[0]), _f(metric[1])
            assert Xm.shape == tuple(n) + (3, 3) and Xj.shape == tuple(n)
            self._metric = (Xm, Xj)  # kept alive for the pointers below
            c.X_metric = Xm.ctypes.data
            c.X_jacobian = Xj.ctypes.data
        self._kvec = None
        if kvector is not None:
            kv = _f(kvector)
            assert kv.shape == tuple(n) + (3,)
            self._kvec = kv
            c.kvector = kv.ctypes.data
        self._penal = None
        if penalization is not None:
            ph, ip = _f(penalization[0]), _f(penalization[1])
            assert ph.shape == tuple(n) and ip.shape == tuple(n)
            self._penal = (ph, ip)
            c.penal_phi = ph.ctypes.data
            c.penal_inv_phi = ip.ctypes.data
        self.c = c

    @property
    def shape(self):
        return tuple(self.c.n)

    def check_adjoint_rows(self):
        """set_boundary_condition_rhss (boundary_conditions.f90:302-377) has
        its cases 204-206 commented out: such rows reach case default, which
        stops the run.  Raised here instead of the C oracle's abort."""
        for r in range(self.c.nbc):
            t = int(round(self.c.bc[r][0]))
            if t in (204, 205, 206):
                raise ValueError(f"boundary type {t} stops set_boundary_condition_rhss")

    def ref(self):
        return C.byref(self.c)


def _f(a):
    return np.asfortranarray(a, dtype=np.float64)


def diff_core(scheme, q, dx, kind=0):
    q = _f(q).copy(order="F")
    if kind == DIFF_O2_EXPLICIT:  # open lines only
        lib().orc_diff_core_o2(q.shape[0], q.size // q.shape[0], q, dx)
        return q
    f = {DIFF_O4_EXPLICIT: lib().orc_diff_core_o4, DIFF_O4_PADE: lib().orc_diff_core_pade}.get(
        kind, lib().orc_diff_core)
    f(scheme, q.shape[0], q.size // q.shape[0], q, dx)
    return q


def filter_core(scheme, q, alpha):
    q = _f(q).copy(order="F")
    lib().orc_filter_core(scheme, q.shape[0], q.size // q.shape[0], q, alpha)
    return q


def diff3d(cfg, d, f, h):
    f = _f(f)
    out = np.zeros_like(f, order="F")
    lib().orc_diff3d(cfg.ref(), d, out, f, h)
    return out


def filter4d(cfg, d, q):
    q = _f(q).copy(order="F")
    nv = 1 if q.ndim == 3 else q.shape[3]
    lib().orc_filter4d(cfg.ref(), d, q, nv)
    return q


def rhs_direct(cfg, q, qref=None, shape=None, heat_unit_spacing=1):
    q = _f(q)
    rhs = np.zeros_like(q, order="F")
    qref = None if qref is None else _f(qref)
    shape = None if shape is None else _f(shape)
    lib().orc_rhs_direct(cfg.ref(), rhs, q, _ptr(qref), _ptr(shape), heat_unit_spacing)
    return rhs


def rhs_adjoint(cfg, qs, q, g=None, shape=None):
    cfg.check_adjoint_rows()
    qs, q = _f(qs), _f(q)
    rhs = np.zeros_like(q, order="F")
    g = None if g is None else _f(g)
    shape = None if shape is None else _f(shape)
    lib().orc_rhs_adjoint(cfg.ref(), rhs, qs, q, _ptr(g), _ptr(shape))
    return rhs


def characteristic_rhs(rhs, q, aim, k, gamma, dt):
    """characteristic_rhs_k1d on m points: rhs, q, aim (m, 5); returns the new rhs."""
    rhs = _f(rhs).copy(order="F")
    m = rhs.shape[0]
    lib().orc_characteristic_rhs_k1d(m, m, rhs, _f(q), _f(aim), _f(np.asarray(k, float)), gamma, dt)
    return rhs


def filter_weighted4d(cfg, d, q, weight=None, strength=None):
    """filter_weighted_iN::filter_4d (o2_e_c_p core) on a copy of q (n1,n2,n3[,nv])."""
    q = _f(q).copy(order="F")
    nv = 1 if q.ndim == 3 else q.shape[3]
    w = None if weight is None else _f(weight)
    s = None if strength is None else _f(strength)
    lib().orc_filter_weighted4d(cfg.ref(), d, q, nv, _ptr(w), _ptr(s))
    return q


def calc_metric(cfg, X, X_calc):
    """grid_metric::calc_metric: X, X_calc (n1,n2,n3,3) -> (X_metric (n1,n2,n3,3,3), X_jacobian)."""
    X, X_calc = _f(X), _f(X_calc)
    Xm = np.zeros(cfg.shape + (3, 3), order="F")
    Xj = np.zeros(cfg.shape, order="F")
    lib().orc_calc_metric(cfg.ref(), X, X_calc, Xm, Xj)
    return Xm, Xj


def rhs_direct_gt(cfg, q, X_metric, X_jacobian, qref=None, shape=None):
    q = _f(q)
    rhs = np.zeros_like(q, order="F")
    qref = None if qref is None else _f(qref)
    shape = None if shape is None else _f(shape)
    lib().orc_rhs_direct_gt(cfg.ref(), rhs, q, _f(X_metric), _f(X_jacobian), _ptr(qref), _ptr(shape))
    return rhs


def rhs_direct_euler_gt(cfg, q, X_metric, X_jacobian, qref=None, shape=None):
    q = _f(q)
    rhs = np.zeros_like(q, order="F")
    qref = None if qref is None else _f(qref)
    shape = None if shape is None else _f(shape)
    lib().orc_rhs_direct_euler_gt(cfg.ref(), rhs, q, _f(X_metric), _f(X_jacobian), _ptr(qref), _ptr(shape))
    return rhs


def time_stepper_direct(cfg, q0, qref=None, shape=None, heat_unit_spacing=1):
    q0 = _f(q0)
    q1 = np.zeros_like(q0, order="F")
    qref = None if qref is None else _f(qref)
    shape = None if shape is None else _f(shape)
    lib().orc_time_stepper_direct(cfg.ref(), q1, q0, _ptr(qref), _ptr(shape), heat_unit_spacing)
    return q1


def time_stepper_adjoint(cfg, qs0, q0, g=None, shape=None):
    cfg.check_adjoint_rows()
    qs0, q0 = _f(qs0), _f(q0)
    qs1 = np.zeros_like(q0, order="F")
    g = None if g is None else _f(g)
    shape = None if shape is None else _f(shape)
    lib().orc_time_stepper_adjoint(cfg.ref(), qs1, qs0, q0, _ptr(g), _ptr(shape))
    return qs1


def run_direct(cfg, q, nsteps, nt0=0, qref=None, shape=None, heat_unit_spacing=1):
    q = _f(q).copy(order="F")
    qref = None if qref is None else _f(qref)
    shape = None if shape is None else _f(shape)
    lib().orc_run_direct(cfg.ref(), q, nt0, nsteps, _ptr(qref), _ptr(shape), heat_unit_spacing)
    return q


def run_direct_conv(cfg, q, nsteps, cfreq, ctol, nt0=0, qref=None, shape=None, heat_unit_spacing=1):
    """cruna_hpc's loop with the convergence check: (q, last residual, step
    at which the tolerance was reached or 0)."""
    q = _f(q).copy(order="F")
    qref = None if qref is None else _f(qref)
    shape = None if shape is None else _f(shape)
    res, ntc = C.c_double(), C.c_int()
    lib().orc_run_direct_conv(cfg.ref(), q, nt0, nsteps, cfreq, ctol, _ptr(qref), _ptr(shape),
                              heat_unit_spacing, C.byref(res), C.byref(ntc))
    return q, res.value, ntc.value


def run_adjoint(cfg, qs, Q, shape=None):
    """Q: direct trajectory, shape (n1, n2, n3, 5, nsteps); Q[..., nt-1] is the state after step nt."""
    cfg.check_adjoint_rows()
    qs = _f(qs).copy(order="F")
    Q = _f(Q)
    shape = None if shape is None else _f(shape)
    lib().orc_run_adjoint(cfg.ref(), qs, Q, Q.shape[4], _ptr(shape))
    return qs


def bc_q(cfg, q, qref=None, shape=None):
    q = _f(q).copy(order="F")
    qref = None if qref is None else _f(qref)
    shape = None if shape is None else _f(shape)
    lib().orc_bc_q(cfg.ref(), q, _ptr(qref), _ptr(shape))
    return q


def bc_rhs(cfg, rhs, q, qref=None, shape=None):
    rhs = _f(rhs).copy(order="F")
    qref = None if qref is None else _f(qref)
    shape = None if shape is None else _f(shape)
    lib().orc_bc_rhs(cfg.ref(), rhs, _f(q), _ptr(qref), _ptr(shape))
    return rhs


def bc_qs(cfg, qs, q, shape=None):
    qs = _f(qs).copy(order="F")
    shape = None if shape is None else _f(shape)
    lib().orc_bc_qs(cfg.ref(), qs, _f(q), _ptr(shape))
    return qs


def characteristic_q(q, aim, k, gamma):
    """q, aim: (npts, 5) arrays; k: outward normal (3,)."""
    q = _f(q).copy(order="F")
    lib().orc_characteristic_q_k1d(q.shape[0], q.shape[0], q, _f(aim), _f(np.asarray(k, float)), gamma)
    return q


def characteristic_qs(qs, q, k, gamma):
    qs = _f(qs).copy(order="F")
    lib().orc_characteristic_qs_k1d(qs.shape[0], qs.shape[0], qs, _f(q), _f(np.asarray(k, float)), gamma)
    return qs


def sponge_shape(cfg, x0, x1, length, amplitude, X1, X2, X3, cos_shape=False):
    shape = np.zeros(cfg.shape, order="F")
    lib().orc_sponge_shape(cfg.ref(), _f(np.asarray(x0, float)), _f(np.asarray(x1, float)),
                           _f(np.asarray(length, float)), amplitude, int(cos_shape),
                           _f(X1), _f(X2), _f(X3), shape)
    return shape


def point_closures(cfg, rho, p):
    L = lib()
    T = np.array([L.orc_T(cfg.ref(), r, pp) for r, pp in zip(rho.ravel(), p.ravel())])
    mu = np.array([L.orc_mu(cfg.ref(), t) for t in T])
    lam = np.array([L.orc_lambda(cfg.ref(), m) for m in mu])
    return T, mu, lam
