"""Python host mirror of the CRUNA hot-path interface over libcruna_hip.so.

Names follow the reference's Fortran surface (cruna-toolkit/public):
  Dx1/Dx2/Dx3                 diff_iN_mpi_cartesian::diff (src/diff/diff_i1_mpi_cartesian.f90:7-12)
  filter_i1/filter_i2/...     filter_iN_mpi_cartesian::filter (src/filter/filter_i1_mpi_cartesian.f90:8-13)
  right_hand_side_direct_3d   src/equations/ns_rho_rhou_p.f90:31
  time_stepper_direct         src/time/rk_o4_e.f90:17
and the device-resident step loop of src/main/cruna_hpc.f90:171-229.
Arrays are numpy float64, Fortran order, shape (n1, n2, n3[, 5]).  Errors
from the library raise CrunaError, as the reference's "error in ..." + stop.
The HIP library is required: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcruna_hip.so")

PERIODIC, OPEN = 0, 1
SUTHERLAND, POWER = 0, 1
FP64, FP32_MIXED = 0, 1
DIFF_O6_COMPACT, DIFF_O4_EXPLICIT, DIFF_O4_PADE, DIFF_O2_EXPLICIT = 0, 1, 2, 3
EQ_NS, EQ_EULER = 0, 1
MAX_BC = 32


class CrunaError(RuntimeError):
    pass


class Cfg(C.Structure):
    """cruna_hip_cfg of include/cruna_hip.h."""
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
        ("heat_unit_spacing", C.c_int),
        ("precision", C.c_int),
        ("tile_lines", C.c_int * 3),
        ("nbc", C.c_int),
        ("bc", (C.c_double * 10) * MAX_BC),
        ("diff_kind", C.c_int * 3),
        ("wfilter_freq", C.c_int * 3),
        ("equations", C.c_int),
    ]


_lib = None


def lib():
    """Load libcruna_hip.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CrunaError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.c_int
        L.cruna_hip_default_cfg.argtypes = [C.POINTER(Cfg)]
        L.cruna_hip_default_cfg.restype = None
        L.cruna_hip_create.argtypes = [C.POINTER(Cfg), C.POINTER(vp)]
        L.cruna_hip_destroy.argtypes = [vp]
        L.cruna_hip_last_error.restype = C.c_char_p
        L.cruna_hip_set_state.argtypes = [vp, dp]
        L.cruna_hip_get_state.argtypes = [vp, dp]
        L.cruna_hip_set_qref.argtypes = [vp, dp]
        L.cruna_hip_set_sponge_shape.argtypes = [vp, dp]
        L.cruna_hip_set_kvector.argtypes = [vp, dp]
        L.cruna_hip_set_metric.argtypes = [vp, dp, dp]
        L.cruna_hip_set_penalization.argtypes = [vp, vp, vp]
        L.cruna_hip_filter_weighted.argtypes = [vp, ip, vp, ip, vp, vp]
        L.cruna_hip_autotune.argtypes = [vp, ip, C.POINTER(C.c_int)]
        L.cruna_hip_ensemble_create.argtypes = [C.POINTER(Cfg), ip, C.POINTER(vp)]
        L.cruna_hip_ensemble_destroy.argtypes = [vp]
        L.cruna_hip_ensemble_member.argtypes = [vp, ip, C.POINTER(vp)]
        L.cruna_hip_ensemble_set_role.argtypes = [vp, ip, ip, ip]
        L.cruna_hip_ensemble_run.argtypes = [vp, ip]
        L.cruna_hip_ensemble_autotune.argtypes = [vp, ip, C.POINTER(C.c_int)]
        L.cruna_hip_ensemble_set_batching.argtypes = [vp, ip]
        L.cruna_hip_ensemble_batched_steps.argtypes = [vp, C.POINTER(C.c_int)]
        L.cruna_hip_diff.argtypes = [vp, ip, dp, dp, C.c_double]
        L.cruna_hip_filter.argtypes = [vp, ip, dp, ip]
        L.cruna_hip_rhs_direct.argtypes = [vp, dp, dp, C.c_double]
        L.cruna_hip_rhs_direct_noalloc.argtypes = [vp, dp, dp, C.c_double]
        L.cruna_hip_rhs_adjoint.argtypes = [vp, dp, dp, dp, dp, C.c_double]
        L.cruna_hip_time_stepper_direct.argtypes = [vp, dp, dp, C.c_double]
        L.cruna_hip_time_stepper_adjoint.argtypes = [vp, dp, dp, dp, dp, C.c_double]
        L.cruna_hip_run_direct.argtypes = [vp, ip, ip]
        L.cruna_hip_synchronize.argtypes = [vp]
        L.cruna_hip_set_adjoint_state.argtypes = [vp, dp]
        L.cruna_hip_get_adjoint_state.argtypes = [vp, dp]
        L.cruna_hip_trajectory_alloc.argtypes = [vp, ip]
        L.cruna_hip_trajectory_set.argtypes = [vp, ip, dp]
        L.cruna_hip_trajectory_get.argtypes = [vp, ip, dp]
        L.cruna_hip_trajectory_record.argtypes = [vp, ip]
        L.cruna_hip_run_adjoint.argtypes = [vp, ip, ip]
        L.cruna_hip_profile_enable.argtypes = [vp, ip]
        L.cruna_hip_profile_read.argtypes = [vp, ip, dp, C.POINTER(C.c_long), C.c_char_p, C.c_size_t]
        L.cruna_hip_profile_count.argtypes = []
        fp = C.POINTER(C.c_float)
        L.cruna_hip_cnn_create.argtypes = [ip, C.POINTER(vp)]
        L.cruna_hip_cnn_destroy.argtypes = [vp]
        L.cruna_hip_cnn_num_params.argtypes = [vp]
        L.cruna_hip_cnn_num_params.restype = C.c_long
        L.cruna_hip_cnn_init_params.argtypes = [vp, C.c_uint]
        L.cruna_hip_cnn_set_params.argtypes = [vp, fp]
        L.cruna_hip_cnn_get_params.argtypes = [vp, fp]
        L.cruna_hip_cnn_set_norm.argtypes = [vp, dp, dp, C.c_double]
        L.cruna_hip_box_filter_state.argtypes = [vp, ip, dp, dp]
        L.cruna_hip_cnn_set_gain.argtypes = [vp, dp]
        L.cruna_hip_subsets_alloc.argtypes = [vp, ip, ip]
        L.cruna_hip_run_direct_subsets.argtypes = [vp, ip, ip]
        L.cruna_hip_run_adjoint_subsets.argtypes = [vp, ip, ip]
        L.cruna_hip_cnn_forward.argtypes = [vp, dp, dp]
        L.cruna_hip_cnn_gradients.argtypes = [vp, dp, dp, fp, fp]
        L.cruna_hip_cnn_add_sample.argtypes = [vp, dp, dp]
        L.cruna_hip_cnn_train.argtypes = [vp, ip, C.c_float, ip, ip, C.c_uint, fp]
        L.cruna_hip_cnn_num_samples.argtypes = [vp]
        L.cruna_hip_set_closure.argtypes = [vp, vp]
        L.cruna_hip_set_convergence.argtypes = [vp, ip, C.c_double]
        L.cruna_hip_get_residual.argtypes = [vp, dp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        for f in ("create", "destroy", "set_state", "get_state", "set_qref", "set_sponge_shape", "set_kvector", "diff",
                  "filter", "rhs_direct", "rhs_direct_noalloc", "rhs_adjoint", "time_stepper_direct", "time_stepper_adjoint",
                  "run_direct", "synchronize", "profile_enable", "profile_read", "profile_count",
                  "set_adjoint_state", "get_adjoint_state", "trajectory_alloc", "trajectory_set",
                  "trajectory_get", "trajectory_record", "run_adjoint", "set_metric", "set_penalization", "filter_weighted", "autotune",
                  "ensemble_create", "ensemble_destroy", "ensemble_member", "ensemble_set_role",
                  "ensemble_run", "ensemble_autotune", "ensemble_set_batching", "ensemble_batched_steps", "cnn_create", "cnn_destroy", "cnn_init_params",
                  "cnn_set_params", "cnn_get_params", "cnn_set_norm", "cnn_forward", "cnn_gradients",
                  "cnn_add_sample", "cnn_train", "set_closure", "box_filter_state", "cnn_set_gain", "subsets_alloc", "run_direct_subsets",
                  "run_adjoint_subsets", "set_convergence", "get_residual"):
            getattr(L, "cruna_hip_" + f).restype = C.c_int
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise CrunaError(lib().cruna_hip_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a, shape):
    a = np.asfortranarray(a, dtype=np.float64)
    if a.shape != shape:
        raise CrunaError(f"array shape {a.shape} does not match the grid {shape}")
    return a


def make_cfg(n, dx, scheme=(PERIODIC,) * 3, filter_alpha=0.35, gamma=1.4, R=1.0, W=1.0, Pr=0.71,
             mu0=1.0 / 1600.0, T0=1.0, visc_law=SUTHERLAND, visc_exp=0.76, dt=1e-3,
             filter_freq=(1, 1, 1), afilter_freq=(0, 0, 0), heat_unit_spacing=1, precision=FP64,
             bc=(), tile_lines=(0, 0, 0), diff_kind=(DIFF_O6_COMPACT,) * 3, wfilter_freq=(0, 0, 0),
             equations=0):
    """cruna_hip_cfg from the reference's parameter names (defaults of the library)."""
    cfg = Cfg()
    lib().cruna_hip_default_cfg(C.byref(cfg))
    for d in range(3):
        cfg.n[d] = int(n[d])
        cfg.dx[d] = float(dx[d])
        cfg.scheme[d] = int(scheme[d])
        cfg.filter_freq[d] = int(filter_freq[d])
        cfg.afilter_freq[d] = int(afilter_freq[d])
        cfg.tile_lines[d] = int(tile_lines[d])
        cfg.diff_kind[d] = int(diff_kind[d])
        cfg.wfilter_freq[d] = int(wfilter_freq[d])
    cfg.filter_alpha = filter_alpha
    cfg.gamma, cfg.R, cfg.W, cfg.Pr, cfg.mu0, cfg.T0 = gamma, R, W, Pr, mu0, T0
    cfg.visc_law, cfg.visc_exp, cfg.dt = visc_law, visc_exp, dt
    cfg.heat_unit_spacing, cfg.precision = heat_unit_spacing, precision
    cfg.equations = int(equations)
    cfg.nbc = len(bc)
    for r, row in enumerate(bc):
        for j, v in enumerate(row):
            cfg.bc[r][j] = float(v)
    return cfg


class Cruna:
    """One device context: grid, schemes, material constants, resident state."""

    def __init__(self, n, dx, _handle=None, **kw):
        cfg = make_cfg(n, dx, **kw)
        self.cfg = cfg
        self.shape3 = tuple(int(v) for v in n)
        self.shape4 = self.shape3 + (5,)
        self._owned = _handle is None
        if _handle is not None:  # a member of an Ensemble, which owns it
            self._h = _handle
            return
        h = C.c_void_p()
        _check(lib().cruna_hip_create(C.byref(cfg), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            if self._owned:
                lib().cruna_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ----- derivative and filter (parity-level entry points) -----
    def Dx(self, d, f, h=None):
        """diff_3d: compact derivative of a 3-D field along direction d; h defaults to 1."""
        f = _f64(f, self.shape3)
        out = np.empty(self.shape3, order="F")
        _check(lib().cruna_hip_diff(self._h, d, _dp(out), _dp(f), 1.0 if h is None else float(h)))
        return out

    def Dx1(self, f, h=None):
        return self.Dx(0, f, h)

    def Dx2(self, f, h=None):
        return self.Dx(1, f, h)

    def Dx3(self, f, h=None):
        return self.Dx(2, f, h)

    def filter(self, d, q):
        """filter_4d along direction d (in place on a copy, returned)."""
        q = np.array(q, dtype=np.float64, order="F", copy=True)
        nv = 1 if q.ndim == 3 else q.shape[3]
        if q.shape[:3] != self.shape3:
            raise CrunaError("field does not match the grid")
        _check(lib().cruna_hip_filter(self._h, d, _dp(q), nv))
        return q

    def filter_weighted(self, d, q, weight=None, strength=None):
        """filter_weighted_iN's filter_4d along direction d with the o2_e_c_p
        core (on a copy, returned); weight/strength None mean 1."""
        q = np.array(q, dtype=np.float64, order="F", copy=True)
        nv = 1 if q.ndim == 3 else q.shape[3]
        if q.shape[:3] != self.shape3:
            raise CrunaError("field does not match the grid")
        w = None if weight is None else _f64(weight, self.shape3)
        s = None if strength is None else _f64(strength, self.shape3)
        _check(lib().cruna_hip_filter_weighted(self._h, d, q.ctypes.data, nv,
                                               None if w is None else w.ctypes.data,
                                               None if s is None else s.ctypes.data))
        return q

    def filter_i1(self, q):
        return self.filter(0, q)

    def filter_i2(self, q):
        return self.filter(1, q)

    def filter_i3(self, q):
        return self.filter(2, q)

    # ----- RHS and RK4 -----
    def right_hand_side_direct_3d(self, q, t=0.0):
        q = _f64(q, self.shape4)
        rhs = np.empty(self.shape4, order="F")
        _check(lib().cruna_hip_rhs_direct(self._h, _dp(rhs), _dp(q), float(t)))
        return rhs

    def right_hand_side_direct_3d_noalloc(self, q, t=0.0):
        """The direct RHS through the allocation-free entry point (static
        device staging; see cruna_hip_rhs_direct_noalloc)."""
        q = _f64(q, self.shape4)
        rhs = np.empty(self.shape4, order="F")
        _check(lib().cruna_hip_rhs_direct_noalloc(self._h, _dp(rhs), _dp(q), float(t)))
        return rhs

    def time_stepper_direct(self, q0, t=0.0):
        q0 = _f64(q0, self.shape4)
        q1 = np.empty(self.shape4, order="F")
        _check(lib().cruna_hip_time_stepper_direct(self._h, _dp(q1), _dp(q0), float(t)))
        return q1

    def right_hand_side_adjoint_3d(self, qs, q, g=None, t=0.0):
        """Adjoint RHS at adjoint state qs about the forward state q; g may be None."""
        qs, q = _f64(qs, self.shape4), _f64(q, self.shape4)
        gp = None if g is None else _dp(_f64(g, self.shape4))
        rhs = np.empty(self.shape4, order="F")
        _check(lib().cruna_hip_rhs_adjoint(self._h, _dp(rhs), _dp(qs), _dp(q), gp, float(t)))
        return rhs

    def time_stepper_adjoint(self, qs0, q0, g=None, t=0.0):
        """One RK4 step backward in time (dt -> -dt), q0 and g frozen over the stages."""
        qs0, q0 = _f64(qs0, self.shape4), _f64(q0, self.shape4)
        g = None if g is None else _f64(g, self.shape4)
        qs1 = np.empty(self.shape4, order="F")
        _check(lib().cruna_hip_time_stepper_adjoint(self._h, _dp(qs1), _dp(qs0), _dp(q0),
                                                    None if g is None else _dp(g), float(t)))
        return qs1

    # ----- device-resident loop -----
    def set_state(self, q):
        q = _f64(q, self.shape4)
        _check(lib().cruna_hip_set_state(self._h, _dp(q)))

    def get_state(self):
        q = np.empty(self.shape4, order="F")
        _check(lib().cruna_hip_get_state(self._h, _dp(q)))
        return q

    def set_qref(self, qref):
        qref = _f64(qref, self.shape4)
        _check(lib().cruna_hip_set_qref(self._h, _dp(qref)))

    def set_sponge_shape(self, shape):
        shape = _f64(shape, self.shape3)
        _check(lib().cruna_hip_set_sponge_shape(self._h, _dp(shape)))

    def set_kvector(self, kvector):
        """Boundary normals per point, kvector (n1,n2,n3,3), for rows 202 / 203 / 205 / 206."""
        k = np.asfortranarray(kvector, dtype=np.float64)
        if k.shape != self.shape3 + (3,):
            raise CrunaError(f"kvector shape {k.shape} does not match {self.shape3 + (3,)}")
        _check(lib().cruna_hip_set_kvector(self._h, _dp(k)))

    def set_metric(self, X_metric, X_jacobian):
        """Deformed grid: X_metric (n1,n2,n3,3,3) and X_jacobian (n1,n2,n3) of calc_metric."""
        Xm = np.asfortranarray(X_metric, dtype=np.float64)
        if Xm.shape != self.shape3 + (3, 3):
            raise CrunaError(f"X_metric shape {Xm.shape} does not match {self.shape3 + (3, 3)}")
        Xj = _f64(X_jacobian, self.shape3)
        _check(lib().cruna_hip_set_metric(self._h, _dp(Xm), _dp(Xj)))

    def set_penalization(self, phi, inv_phi=None):
        """Volume penalisation (ns_rho_rhou_p_ev_phi): porosity phi (n1,n2,n3) and
        the stored 1/phi (default 1/phi); phi None removes it."""
        if phi is None:
            _check(lib().cruna_hip_set_penalization(self._h, None, None))
            return
        ph = _f64(phi, self.shape3)
        ip = _f64(1.0 / ph if inv_phi is None else inv_phi, self.shape3)
        _check(lib().cruna_hip_set_penalization(self._h, ph.ctypes.data, ip.ctypes.data))

    def autotune(self, steps=2):
        """Pick the lines per work-group of each direction by timing; returns them."""
        out = (C.c_int * 3)()
        _check(lib().cruna_hip_autotune(self._h, int(steps), out))
        for d in range(3):
            self.cfg.tile_lines[d] = out[d]
        return tuple(out)

    def run_direct(self, nsteps, nt0=0):
        _check(lib().cruna_hip_run_direct(self._h, int(nt0), int(nsteps)))

    def synchronize(self):
        _check(lib().cruna_hip_synchronize(self._h))

    def set_convergence(self, cfreq, ctol):
        """run.cfreq / run.ctol of cruna_hpc: stop run_direct once
        max|qn - q1|/dt <= ctol at a step with mod(nt, cfreq) == 0 (0: off)."""
        _check(lib().cruna_hip_set_convergence(self._h, int(cfreq), float(ctol)))

    def residual(self):
        """(last residual or -1, step it was computed at, step at which the
        last run_direct reached the tolerance or 0)."""
        r, ntc, ntv = C.c_double(), C.c_int(), C.c_int()
        _check(lib().cruna_hip_get_residual(self._h, C.byref(r), C.byref(ntc), C.byref(ntv)))
        return r.value, ntc.value, ntv.value

    # ----- adjoint loop -----
    def set_adjoint_state(self, qs):
        qs = _f64(qs, self.shape4)
        _check(lib().cruna_hip_set_adjoint_state(self._h, _dp(qs)))

    def get_adjoint_state(self):
        qs = np.empty(self.shape4, order="F")
        _check(lib().cruna_hip_get_adjoint_state(self._h, _dp(qs)))
        return qs

    def trajectory_alloc(self, nslots):
        """Device slots for Q(:,:,:,:,nt), nt = 1..nslots (state after forward step nt)."""
        _check(lib().cruna_hip_trajectory_alloc(self._h, int(nslots)))

    def trajectory_set(self, nt, q):
        q = _f64(q, self.shape4)
        _check(lib().cruna_hip_trajectory_set(self._h, int(nt), _dp(q)))

    def trajectory_get(self, nt):
        q = np.empty(self.shape4, order="F")
        _check(lib().cruna_hip_trajectory_get(self._h, int(nt), _dp(q)))
        return q

    def trajectory_record(self, on=True):
        _check(lib().cruna_hip_trajectory_record(self._h, 1 if on else 0))

    def run_adjoint(self, nsteps, nt_end=None):
        """calc_adjoint's loop for nt = nt_end .. nt_end-nsteps+1 (nt_end defaults to nsteps)."""
        nt_end = nsteps if nt_end is None else nt_end
        _check(lib().cruna_hip_run_adjoint(self._h, int(nt_end), int(nsteps)))

    def box_filter_state(self, f, rhs=True):
        """(q_box, rhs_box or None): f^3 box averages of the device state and of its direct RHS."""
        shape = tuple(n // f for n in self.shape3) + (5,)
        qb = np.zeros(shape, order="F")
        rb = np.zeros(shape, order="F") if rhs else None
        _check(lib().cruna_hip_box_filter_state(self._h, int(f), _dp(qb), None if rb is None else _dp(rb)))
        return qb, rb

    def subsets_alloc(self, subsets, steps):
        """Checkpoints for `subsets` subsets and a `steps`-slot trajectory."""
        _check(lib().cruna_hip_subsets_alloc(self._h, int(subsets), int(steps)))

    def run_direct_subsets(self, subsets, steps):
        """calc_direct over subsets (nt = 1..steps in each), checkpointing every subset start."""
        _check(lib().cruna_hip_run_direct_subsets(self._h, int(subsets), int(steps)))

    def run_adjoint_subsets(self, subsets, steps):
        """calc_adjoint over subsets, last first, recomputing each subset's trajectory."""
        _check(lib().cruna_hip_run_adjoint_subsets(self._h, int(subsets), int(steps)))

    def set_closure(self, closure):
        """Add a learned closure (a `Closure` of this grid) to the direct RHS; None removes it."""
        _check(lib().cruna_hip_set_closure(self._h, None if closure is None else closure._h))
        self._closure = closure

    def profile(self, on=True):
        _check(lib().cruna_hip_profile_enable(self._h, 1 if on else 0))

    def profile_read(self):
        out = {}
        for k in range(lib().cruna_hip_profile_count()):
            ms = C.c_double()
            n = C.c_long()
            name = C.create_string_buffer(64)
            _check(lib().cruna_hip_profile_read(self._h, k, C.byref(ms), C.byref(n), name, 64))
            if n.value:
                out[name.value.decode()] = (ms.value, n.value)
        return out


FORWARD, ADJOINT = 0, 1


class Ensemble:
    """Independent members (dicts of Cruna keyword arguments with n and dx) on
    one device; run() advances all of them through a replayed HIP graph."""

    def __init__(self, members):
        cfgs = (Cfg * len(members))()
        for i, kw in enumerate(members):
            kw = dict(kw)
            cfgs[i] = make_cfg(kw.pop("n"), kw.pop("dx"), **kw)
        h = C.c_void_p()
        _check(lib().cruna_hip_ensemble_create(cfgs, len(members), C.byref(h)))
        self._h = h
        self.members = []
        for i, kw in enumerate(members):
            mh = C.c_void_p()
            _check(lib().cruna_hip_ensemble_member(h, i, C.byref(mh)))
            kw = dict(kw)
            self.members.append(Cruna(kw.pop("n"), kw.pop("dx"), _handle=mh, **kw))

    def __len__(self):
        return len(self.members)

    def __getitem__(self, i):
        return self.members[i]

    def set_role(self, i, role, nt):
        _check(lib().cruna_hip_ensemble_set_role(self._h, int(i), int(role), int(nt)))

    def run(self, nsteps):
        _check(lib().cruna_hip_ensemble_run(self._h, int(nsteps)))

    def set_batching(self, bins=-1):
        """Batched launches when the members allow them: `bins` concurrent
        batches (-1 automatic, the default; 1: one launch per sweep kind over
        all members); 0 or False keeps one replayed graph per member."""
        if bins is True:
            bins = -1
        _check(lib().cruna_hip_ensemble_set_batching(self._h, int(bins)))

    def batched_steps(self):
        """Steps of the last run() that went through batched launches."""
        n = C.c_int()
        _check(lib().cruna_hip_ensemble_batched_steps(self._h, C.byref(n)))
        return n.value

    def autotune(self, steps=2):
        """Tile autotuner on every member; returns the lines per work-group per member."""
        out = (C.c_int * (3 * len(self.members)))()
        _check(lib().cruna_hip_ensemble_autotune(self._h, int(steps), out))
        res = []
        for i, m in enumerate(self.members):
            t = tuple(out[3 * i + d] for d in range(3))
            for d in range(3):
                m.cfg.tile_lines[d] = t[d]
            res.append(t)
        return res

    def close(self):
        if getattr(self, "_h", None):
            for m in self.members:
                m._h = None
            lib().cruna_hip_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Closure:
    """Learned subfilter closure: the 3-D CNN of cruna_hip_cnn on a periodic
    S^3 grid (trained and evaluated on the device).  Fields are (S,S,S,5)
    float64 arrays in Fortran order like the state."""

    def __init__(self, S, seed=None):
        h = C.c_void_p()
        _check(lib().cruna_hip_cnn_create(int(S), C.byref(h)))
        self._h = h
        self.S = int(S)
        self.shape = (self.S,) * 3 + (5,)
        if seed is not None:
            _check(lib().cruna_hip_cnn_init_params(h, int(seed)))

    def num_params(self):
        return int(lib().cruna_hip_cnn_num_params(self._h))

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != self.num_params():
            raise CrunaError(f"expected {self.num_params()} parameters, got {p.size}")
        _check(lib().cruna_hip_cnn_set_params(self._h, _fp(p)))

    def get_params(self):
        p = np.zeros(self.num_params(), dtype=np.float32)
        _check(lib().cruna_hip_cnn_get_params(self._h, _fp(p)))
        return p

    def set_norm(self, in_std, out_std, p_ref=0.0):
        """Input scales (about the sample's field means), output scales, and
        the reference mean pressure of the pressure output (0: unscaled)."""
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (in_std, out_std)]
        if any(x.shape != (5,) for x in a):
            raise CrunaError("normalisation needs 5 values per argument")
        _check(lib().cruna_hip_cnn_set_norm(self._h, _dp(a[0]), _dp(a[1]), float(p_ref)))

    def set_gain(self, gain):
        """Scale of the closure term at inference, one value or one per
        variable (training is unaffected)."""
        g = np.ascontiguousarray(np.broadcast_to(np.asarray(gain, dtype=np.float64), (5,)))
        _check(lib().cruna_hip_cnn_set_gain(self._h, _dp(g)))

    def forward(self, q):
        q = _f64(q, self.shape)
        out = np.zeros(self.shape, order="F")
        _check(lib().cruna_hip_cnn_forward(self._h, _dp(q), _dp(out)))
        return out

    def gradients(self, q, target):
        """(loss, gradient of every parameter) for one (state, target) pair."""
        q, t = _f64(q, self.shape), _f64(target, self.shape)
        g = np.zeros(self.num_params(), dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        _check(lib().cruna_hip_cnn_gradients(self._h, _dp(q), _dp(t), _fp(g), _fp(loss)))
        return float(loss[0]), g

    def add_sample(self, q, target):
        q, t = _f64(q, self.shape), _f64(target, self.shape)
        _check(lib().cruna_hip_cnn_add_sample(self._h, _dp(q), _dp(t)))

    def num_samples(self):
        return int(lib().cruna_hip_cnn_num_samples(self._h))

    def train(self, iters, lr=1e-3, batch=8, crop=32, seed=0):
        """Adam steps on batches of `batch` crop^3 windows of the stored
        samples (crop 0: whole samples); returns the loss of every step."""
        losses = np.zeros(int(iters), dtype=np.float32)
        _check(lib().cruna_hip_cnn_train(self._h, int(iters), float(lr), int(batch), int(crop), int(seed),
                                         _fp(losses)))
        return losses

    def close(self):
        """Free the closure; CrunaError while a context still uses it."""
        if getattr(self, "_h", None):
            _check(lib().cruna_hip_cnn_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
