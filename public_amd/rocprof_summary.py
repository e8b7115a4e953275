"""Reduce rocprofv3 result databases of a libcruna_hip run to per-kernel tables.

    python3 -m public_amd.rocprof_summary trace  <kernel-trace results.db>
    python3 -m public_amd.rocprof_summary traffic <FETCH_SIZE results.db> <WRITE_SIZE results.db> [copy_bytes]

`trace` lists every kernel with its dispatch count and average duration.
`traffic` lists FETCH_SIZE and WRITE_SIZE per dispatch (KiB, as the counters
report them) and the HBM bytes per dispatch after scaling each counter by what
it reads for the calibration copies in the same run: bench.py
--copy-calibration issues three device-to-device copies of 1 GiB each, which
are the run's largest copy dispatches (other copies of the run, e.g. state
copies of a 256^3 box, are smaller); the scale makes each of those copies read
and write `copy_bytes` (default 1 GiB).

Library kernels are named as the library's profiler names them (sweep1_x,
sweep2_y_rk, filter_z, asweep_x_rk, ...) with a /fp64 or /fp32 suffix for the
arithmetic type; other kernels keep a shortened symbol.
"""
import collections
import re
import sqlite3
import sys

_SWEEP_NAMES = {
    ("k_sweep1", "0"): "sweep1_x", ("k_sweep1", "1"): "sweep1_y", ("k_sweep1", "2"): "sweep1_z",
    ("k_sweep2", "2"): "sweep2_z", ("k_sweep2", "1"): "sweep2_y", ("k_sweep2", "0"): "sweep2_x_rk",
    ("k_filter", "0"): "filter_x", ("k_filter", "1"): "filter_y", ("k_filter", "2"): "filter_z",
    ("k_asweep", "0"): "asweep_x_rk", ("k_asweep", "1"): "asweep_y", ("k_asweep", "2"): "asweep_z",
    ("k_diff", "0"): "diff", ("k_diff", "1"): "diff", ("k_diff", "2"): "diff",
}


def kernel_name(symbol):
    """Library name of a kernel symbol (demangled or not)."""
    m = re.search(r"cruna::(k_\w+)<([^>]*)>", symbol)
    if m:
        args = [a.strip() for a in m.group(2).split(",")]
        kind = "fp64" if args[0] == "double" else "fp32"
        d = next((a for a in args if a in ("0", "1", "2")), None)
        base = _SWEEP_NAMES.get((m.group(1), d), m.group(1) + "<" + ",".join(args) + ">")
        if m.group(1) in ("k_sweep12_z", "kb_sweep12_z"):
            base = "sweep12_z"
        return f"{base}/{kind}"
    if "copyBuffer" in symbol or "copy" in symbol.lower():
        return "copy:" + symbol[:40]
    return symbol[:48]


def _rows(db, query):
    con = sqlite3.connect(db)
    try:
        return con.execute(query).fetchall()
    finally:
        con.close()


def trace(db):
    """Per (kernel, work-group count): dispatches, summed duration, grid, wg.
    The same kernel on grids of other sizes (e.g. 64^3 and 256^3) gets a row
    of its own."""
    agg = collections.OrderedDict()
    for name, dur, gx, wx in _rows(db, "select name, duration, grid_x, workgroup_x from kernels"):
        nb = gx // max(wx, 1)
        a = agg.setdefault((kernel_name(name), nb), [0, 0.0, nb, wx])
        a[0] += 1
        a[1] += dur
    return agg


def counters(db, counter):
    agg = collections.OrderedDict()
    q = "select kernel_name, counter_name, value from counters_collection"
    for name, cn, val in _rows(db, q):
        if cn != counter:
            continue
        a = agg.setdefault(kernel_name(name), [0, 0.0])
        a[0] += 1
        a[1] += val
    return agg


def calibration_copies(db, counter):
    """Counter values (KiB) of the calibration copies: the copy dispatches
    within 1 % of the largest copy dispatch."""
    vals = [v for name, cn, v in _rows(db, "select kernel_name, counter_name, value from counters_collection")
            if cn == counter and kernel_name(name).startswith("copy:")]
    if not vals:
        return []
    top = max(vals)
    return [v for v in vals if v >= 0.99 * top]


def print_trace(db):
    agg = trace(db)
    print(f"{'kernel':34s} {'dispatches':>10s} {'avg_us':>10s} {'grid':>8s} {'wg':>5s}")
    for (name, _), (n, tot, g, w) in agg.items():
        print(f"{name:34s} {n:10d} {tot / n / 1e3:10.2f} {g:8d} {w:5d}")


def print_traffic(fetch_db, write_db, copy_bytes=float(1 << 30)):
    fe, wr = counters(fetch_db, "FETCH_SIZE"), counters(write_db, "WRITE_SIZE")
    cf, cw = calibration_copies(fetch_db, "FETCH_SIZE"), calibration_copies(write_db, "WRITE_SIZE")
    fscale = wscale = None
    if cf and cw:
        fscale = len(cf) * copy_bytes / (sum(cf) * 1024.0)
        wscale = len(cw) * copy_bytes / (sum(cw) * 1024.0)
        print(f"calibration: {len(cf)} / {len(cw)} copies of {copy_bytes:.0f} B read FETCH_SIZE {sum(cf):.0f} KiB and "
              f"WRITE_SIZE {sum(cw):.0f} KiB -> scale fetch x{fscale:.3f}, write x{wscale:.3f}")
    print(f"{'kernel':34s} {'FETCH_KiB':>12s} {'WRITE_KiB':>12s} {'hbm_bytes':>14s}")
    for name, (n, v) in fe.items():
        f = v / n
        w = wr.get(name, (1, 0.0))
        w = w[1] / max(w[0], 1)
        b = "" if fscale is None else f"{f * 1024 * fscale + w * 1024 * wscale:14.0f}"
        print(f"{name:34s} {f:12.0f} {w:12.0f} {b}")


if __name__ == "__main__":
    if len(sys.argv) >= 3 and sys.argv[1] == "trace":
        print_trace(sys.argv[2])
    elif len(sys.argv) >= 4 and sys.argv[1] == "traffic":
        print_traffic(sys.argv[2], sys.argv[3], float(sys.argv[4]) if len(sys.argv) > 4 else float(1 << 30))
    else:
        print(__doc__)
        sys.exit(2)
