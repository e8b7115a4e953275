"""Snapshot / restart files (public_amd/snapshot.py over libcruna_io.so)
against the format of src/io/hdf5_single.f90 and the names of
src/helper/io_helper.f90; the file structure is read back with HDF5's own
h5dump, a reader independent of the library."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "public_amd", "libcruna_io.so")),
                                reason="libcruna_io.so not built (no HDF5)")


def _io():
    from public_amd import snapshot
    return snapshot


def test_file_names_follow_get_fname():
    s = _io()
    # fn_general "data_direct_snapshot_" + "_" block I2.2 + "_" image I5.5 + "_" ns I2.2 + "_" nt I7.7
    assert s.fname("data_direct_snapshot", "/w", 1, 1, 2, 150) == "/w/data_direct_snapshot__01_00001_02_0000150.h5"
    assert s.fname("data_adjoint_snapshot", ".", 3, 12, 1, 0) == "./data_adjoint_snapshot__03_00012_01_0000000.h5"
    assert s.fname("data_direct", "run", 1, 7, 4) == "run/data_direct__01_00007_04.h5"
    assert s.fname("restartfile", "/w") == "/w/restartfile.h5"
    with pytest.raises(s.CrunaIOError, match="not one of"):
        s.fname("ballon_loop")


def test_state_round_trip_is_exact(tmp_path):
    s = _io()
    rng = np.random.default_rng(0)
    q = np.asfortranarray(rng.standard_normal((16, 12, 8, 5)))
    path = str(tmp_path / "a.h5")
    s.store_4d(path, q, ["geom.n1=16", "time.nt=3"])
    back = s.load_4d(path, q.shape)
    assert np.array_equal(back, q)
    assert s.read_params(path) == ["geom.n1=16", "time.nt=3"]
    with pytest.raises(s.CrunaIOError, match="dimensions"):
        s.load_4d(path, (16, 12, 8, 4))
    with pytest.raises(s.CrunaIOError, match="cannot open"):
        s.load_4d(str(tmp_path / "missing.h5"), q.shape)


def test_restart_round_trip(tmp_path):
    s = _io()
    path = str(tmp_path / "restartfile.h5")
    s.store_restart(path, 3, 1234, ["time.steps=2000"])
    assert s.load_restart(path) == (3, 1234)


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not found")
def test_layout_read_back_by_h5dump(tmp_path):
    """The datasets as hdf5_single.f90 writes them: "data" native double with
    the Fortran dimensions reversed in HDF5's C order, element (i,j,k,v) at
    the Fortran position; "params" 512 fixed strings of 1920 blank-padded
    characters (h5t_fortran_s1)."""
    s = _io()
    q = np.zeros((3, 2, 2, 2), order="F")
    q[..., 0] = np.arange(12).reshape((3, 2, 2), order="F")
    q[..., 1] = 100 + np.arange(12).reshape((3, 2, 2), order="F")
    path = str(tmp_path / "b.h5")
    s.store_4d(path, q, ["geom.n1=3;4"])
    head = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, check=True).stdout
    assert 'DATASET "data"' in head and "H5T_IEEE_F64LE" in head
    assert "DATASPACE  SIMPLE { ( 2, 2, 2, 3 ) / ( 2, 2, 2, 3 ) }" in head
    assert 'DATASET "params"' in head and "STRSIZE 1920" in head and "H5T_STR_SPACEPAD" in head
    assert "DATASPACE  SIMPLE { ( 512 ) / ( 512 ) }" in head
    vals = subprocess.run([H5DUMP, "-d", "data", "-y", "-w", "0", path], capture_output=True, text=True,
                          check=True).stdout
    body = vals[vals.index("DATA {") + 6:vals.index("}", vals.index("DATA {"))]
    nums = [float(x) for x in body.replace(",", " ").split()]
    assert nums == list(np.arange(12.0)) + list(100 + np.arange(12.0))
    rows = subprocess.run([H5DUMP, "-d", "params", "-y", "-w", "0", path], capture_output=True,
                          text=True, check=True).stdout
    assert '"geom.n1=3;4' in rows
