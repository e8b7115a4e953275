"""CPU-side checks of the C-ABI library: it loads and exports every entry
point declared in include/cruna_hip.h (no device calls are made)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cruna_hip.h")
LIB = os.path.join(ROOT, "public_amd", "libcruna_hip.so")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(cruna_hip_[a-z_0-9]+)\s*\(", text)))


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    syms = declared_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_default_cfg_and_errors_without_device():
    from public_amd.cruna import Cfg, lib
    L = lib()
    cfg = Cfg()
    L.cruna_hip_default_cfg(ctypes.byref(cfg))
    assert abs(cfg.filter_alpha - 0.35) < 1e-15 and abs(cfg.Pr - 0.71) < 1e-15
    assert cfg.heat_unit_spacing == 1
    # invalid configuration is rejected before any device work
    cfg.n[0] = 20
    h = ctypes.c_void_p()
    rc = L.cruna_hip_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc != 0 and b"multiple of 8" in L.cruna_hip_last_error()


def test_cfg_struct_matches_header():
    """The ctypes mirror of cruna_hip_cfg lists the header's fields in order."""
    from public_amd.cruna import Cfg
    text = open(HEADER).read()
    body = text[text.index("typedef struct cruna_hip_cfg {"):text.index("} cruna_hip_cfg;")]
    names = re.findall(r"\b(?:int|double)\s+([a-zA-Z_0-9]+)", body)
    assert names == [f[0] for f in Cfg._fields_]


IO_HEADER = os.path.join(ROOT, "include", "cruna_io.h")
IO_LIB = os.path.join(ROOT, "public_amd", "libcruna_io.so")


@pytest.mark.skipif(not os.path.exists(IO_LIB), reason="IO library not built (no HDF5)")
def test_io_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(IO_LIB)
    syms = sorted(set(re.findall(r"\b(cruna_io_[a-z_0-9]+)\s*\(", open(IO_HEADER).read())))
    assert len(syms) >= 8
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
