"""CPU check of the device line-solve algorithm: the host tables of
public_amd/csrc/line_ops.hpp, applied by a CPU emulation of the block engine,
reproduce the oracle's LU-based derivative (o6 compact, o4 explicit, o4 Pade, o2 explicit)
and filter cores to rounding."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_block_engine_matches_lu_cores(tmp_path):
    exe = tmp_path / "line_ops_check"
    src = os.path.join(ROOT, "tests", "cpp", "line_ops_check.cpp")
    orc = os.path.join(ROOT, "oracle", "cruna_oracle.c")
    obj = tmp_path / "oracle.o"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-c", orc, "-o", str(obj)], check=True)
    subprocess.run(["g++", "-O2", "-std=c++17", src, str(obj), "-o", str(exe), "-lm"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    rows = [line.split() for line in out.strip().splitlines()]
    assert len(rows) == 72 + 5 * 10
    for kind, scheme, n, K, err in rows:
        assert float(err) < 1e-14, (kind, scheme, n, K, err)
