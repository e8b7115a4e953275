"""Run calls into the reference's compiled leaf modules in a fresh process.

The reference cores keep their factorisation in `save` variables bound to the
first line length they see (src/diff/diff_pen_core_o6_i_c_p.f90:27-32), so
each configuration gets its own process.
"""
import multiprocessing as mp
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _worker(fn_name, args, q):
    sys.path.insert(0, ROOT)
    try:
        import tests.refproc as rp
        q.put(("ok", getattr(rp, fn_name)(*args)))
    except Exception as e:  # pragma: no cover - reported to the parent
        q.put(("err", repr(e)))


def run(fn_name, *args):
    ctx = mp.get_context("spawn")
    q = ctx.Queue()
    p = ctx.Process(target=_worker, args=(fn_name, args, q))
    p.start()
    status, val = q.get(timeout=120)
    p.join(timeout=30)
    if status != "ok":
        raise RuntimeError(val)
    return val


def available():
    return os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libcruna_ref.so"))


def _lib(alpha=0.35, gamma=1.4, R=1.0, W=1.0, Pr=0.71, mu0=1e-3, T0=1.0, npts=1):
    from oracle import oracle as O
    L = O.ref_lib()
    L.ref_init(alpha, gamma, R, W, Pr, mu0, T0, npts)
    return L


def core(kind, scheme, q, arg):
    """kind 'diff', 'diff_o4' or 'diff_pade' (arg = dx) or 'filter' (arg = alpha) on q(n, nlines)."""
    import numpy as np
    q = np.asfortranarray(q, dtype=np.float64).copy(order="F")
    if kind == "diff":
        L = _lib()
        L.ref_diff(scheme, q, q.shape[0], q.shape[1], arg)
    elif kind == "diff_o4":
        L = _lib()
        L.ref_diff_o4(scheme, q, q.shape[0], q.shape[1], arg)
    elif kind == "diff_pade":
        L = _lib()
        L.ref_diff_pade(scheme, q, q.shape[0], q.shape[1], arg)
    elif kind == "diff_o2":
        L = _lib()
        L.ref_diff_o2(q, q.shape[0], q.shape[1], arg)
    else:
        L = _lib(alpha=arg)
        L.ref_filter(scheme, q, q.shape[0], q.shape[1])
    return q


def trafos(q, gamma, R, W, Pr, mu0, T0):
    import numpy as np
    n = q.shape[0]
    L = _lib(gamma=gamma, R=R, W=W, Pr=Pr, mu0=mu0, T0=T0, npts=n)
    T = np.zeros(n, order="F"); mu = np.zeros(n, order="F"); lam = np.zeros(n, order="F")
    L.ref_trafos(np.asfortranarray(q), n, T, mu, lam)
    return T, mu, lam


def characteristic(kind, y, other, k, gamma):
    import numpy as np
    n = y.shape[0]
    L = _lib(gamma=gamma, npts=n)
    y = np.asfortranarray(y, dtype=np.float64).copy(order="F")
    kk = np.asfortranarray(np.asarray(k, dtype=np.float64))
    if kind == "q":
        L.ref_characteristic_q(y, np.asfortranarray(other), n, kk)
    else:
        L.ref_characteristic_qs(y, np.asfortranarray(other), n, kk)
    return y


def sponge(kind, y, shape, q, qref, gamma):
    import numpy as np
    n = y.shape[0]
    L = _lib(gamma=gamma, npts=n)
    y = np.asfortranarray(y, dtype=np.float64).copy(order="F")
    L.ref_sponge(kind, y, np.asfortranarray(shape), np.asfortranarray(q), np.asfortranarray(qref), n)
    return y


def step_with_reference_cores(q, h, kw, threads):
    """One step of the oracle's loop with the reference's compiled o6 cores
    (serial, or its OpenMP derivative core with `threads` threads)."""
    from oracle import oracle as O
    n = q.shape[0]
    if not O.use_reference_cores(threads=threads):
        raise RuntimeError("oracle/_ref is not built")
    cfg = O.Cfg((n,) * 3, (h,) * 3, **kw)
    return O.run_direct(cfg, q, 1)
