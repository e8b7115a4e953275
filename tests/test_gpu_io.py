"""A device run that stores a snapshot and restart file the way cruna_hpc
does and resumes from them in a fresh context continues bit for bit."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "public_amd", "libcruna_io.so")),
                    reason="libcruna_io.so not built (no HDF5)")
def test_snapshot_resume_continues_exactly(tmp_path):
    from public_amd import Cruna, ic, snapshot
    n, mach = 32, 0.3
    h = ic.periodic_dx(n)
    q0 = ic.periodic_box_state(n, 11, mach)
    kw = dict(mu0=1.0 / 1600.0, T0=1.0 / (1.4 * mach * mach), dt=ic.cfl_dt(q0, h), filter_freq=(1, 2, 3))
    with Cruna((n,) * 3, (h,) * 3, **kw) as g:
        g.set_state(q0)
        g.run_direct(6)
        g.synchronize()
        straight = g.get_state()
    wd = str(tmp_path)
    with Cruna((n,) * 3, (h,) * 3, **kw) as g:
        g.set_state(q0)
        g.run_direct(3)
        g.synchronize()
        path = snapshot.store_snapshot(g, wd, ns=1, nt=3)
    assert os.path.basename(path) == "data_direct_snapshot__01_00001_01_0000003.h5"
    with Cruna((n,) * 3, (h,) * 3, **kw) as g:
        ns, nt = snapshot.resume(g, wd)
        assert (ns, nt) == (1, 3)
        g.run_direct(3, nt0=nt)
        g.synchronize()
        resumed = g.get_state()
    assert np.array_equal(resumed, straight)
    rows = snapshot.read_params(path)
    assert "geom.n1=32" in rows and "time.nt=3" in rows
