"""Snapshot and restart files in CRUNA's format (libcruna_io.so, HDF5).

A device run writes what cruna_hpc writes at its snapshot times
(src/main/cruna_hpc.f90:100-103, 250-262): the state as
data_direct_snapshot__<block>_<image>_<ns>_<nt>.h5 with datasets "data"
(n1, n2, n3, 5 in Fortran order) and "params", and the restart file
(ns, nt); and it resumes from them as the restart procedure does
(:111-148).  The files are those of src/io/hdf5_single.f90, so the
reference's own reader (ext/python/read_cruna_data.py) and cruna_hpc read
them too.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NPARAMS = 512
PARAM_LEN = 384 * 5

_lib = None


class CrunaIOError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "libcruna_io.so")
        if not os.path.exists(path):
            raise CrunaIOError("public_amd/libcruna_io.so is not built (needs HDF5; run __graft_entry__.build())")
        L = C.CDLL(path)
        vp, ip = C.c_void_p, C.c_int
        L.cruna_io_fname.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p, C.c_char_p, ip, ip, ip, ip]
        L.cruna_io_store_4d.argtypes = [C.c_char_p, vp, vp, vp, ip]
        L.cruna_io_load_4d.argtypes = [C.c_char_p, vp, vp]
        L.cruna_io_store_restart.argtypes = [C.c_char_p, ip, ip, vp, ip]
        L.cruna_io_load_restart.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.cruna_io_param_count.argtypes = [C.c_char_p, C.POINTER(C.c_int)]
        L.cruna_io_param_row.argtypes = [C.c_char_p, ip, C.c_char_p, C.c_size_t]
        L.cruna_io_last_error.restype = C.c_char_p
        for f in ("fname", "store_4d", "load_4d", "store_restart", "load_restart", "param_count",
                  "param_row"):
            getattr(L, "cruna_io_" + f).restype = C.c_int
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise CrunaIOError(lib().cruna_io_last_error().decode())


def _rows(params):
    rows = [r.encode() for r in (params or [])]
    arr = (C.c_char_p * max(1, len(rows)))(*rows) if rows else None
    return arr, len(rows)


def fname(type, workdir=".", world_block=1, block_image=1, ns=1, nt=0):
    """get_fname + ".h5" for the snapshot, subset and restart types."""
    buf = C.create_string_buffer(4096)
    _check(lib().cruna_io_fname(buf, 4096, type.encode(), workdir.encode(), int(world_block),
                                int(block_image), int(ns), int(nt)))
    return buf.value.decode()


def param_rows(cfg, ns, nt, extra=None):
    """create_parameter_list rows ("name=v1;v2;...") for a context's cfg: the
    geometry, time, material and filter settings a reader of the files
    needs (read_cruna_data.py takes geom.n1..n3 and parallelism.i1..i3)."""
    rows = [f"geom.n{d + 1}={cfg.n[d]}" for d in range(3)]
    rows += [f"geom.dx{d + 1}={cfg.dx[d]!r}" for d in range(3)]
    rows += [f"parallelism.i{d + 1}=1" for d in range(3)]
    rows += [f"time.ns={int(ns)}", f"time.nt={int(nt)}", f"time.dt={cfg.dt!r}"]
    rows += [f"material.gamma={cfg.gamma!r}", f"material.R={cfg.R!r}", f"material.W={cfg.W!r}",
             f"material.Pr={cfg.Pr!r}", f"material.mu0={cfg.mu0!r}", f"material.T0={cfg.T0!r}"]
    rows += [f"filter.alpha={cfg.filter_alpha!r}"]
    rows += [f"filter.di{d + 1}freq={cfg.filter_freq[d]}" for d in range(3)]
    for k, v in (extra or {}).items():
        vals = v if isinstance(v, (list, tuple)) else [v]
        rows.append(f"{k}=" + ";".join(str(x) for x in vals))
    return rows


def store_4d(path, data, params=None):
    a = np.asfortranarray(data, dtype=np.float64)
    if a.ndim == 3:
        a = a[..., None]
    dims = (C.c_int * 4)(*a.shape)
    rows, n = _rows(params)
    _check(lib().cruna_io_store_4d(path.encode(), a.ctypes.data, dims, rows, n))


def load_4d(path, shape):
    a = np.empty(tuple(shape), order="F")
    dims = (C.c_int * 4)(*shape)
    _check(lib().cruna_io_load_4d(path.encode(), a.ctypes.data, dims))
    return a


def store_restart(path, ns, nt, params=None):
    rows, n = _rows(params)
    _check(lib().cruna_io_store_restart(path.encode(), int(ns), int(nt), rows, n))


def load_restart(path):
    ns, nt = C.c_int(), C.c_int()
    _check(lib().cruna_io_load_restart(path.encode(), C.byref(ns), C.byref(nt)))
    return ns.value, nt.value


def read_params(path):
    """The non-blank rows of "params"."""
    n = C.c_int()
    _check(lib().cruna_io_param_count(path.encode(), C.byref(n)))
    out = []
    buf = C.create_string_buffer(PARAM_LEN + 1)
    for i in range(NPARAMS):
        _check(lib().cruna_io_param_row(path.encode(), i, buf, PARAM_LEN + 1))
        if buf.value:
            out.append(buf.value.decode())
        if len(out) == n.value:
            break
    return out


def store_snapshot(ctx, workdir, ns, nt, restart=True, extra=None):
    """cruna_hpc's snapshot of the device state (data_direct_snapshot at
    (ns, nt)) and, with restart, the restart file pointing at it."""
    path = fname("data_direct_snapshot", workdir, ns=ns, nt=nt)
    rows = param_rows(ctx.cfg, ns, nt, extra)
    store_4d(path, ctx.get_state(), rows)
    if restart:
        store_restart(fname("restartfile", workdir), ns, nt, rows)
    return path


def resume(ctx, workdir):
    """The restart procedure: (ns, nt) from the restart file, the state of
    that snapshot into the device; returns (ns, nt)."""
    ns, nt = load_restart(fname("restartfile", workdir))
    ctx.set_state(load_4d(fname("data_direct_snapshot", workdir, ns=ns, nt=nt), ctx.shape4))
    return ns, nt
