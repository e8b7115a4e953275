"""MI355X (gfx950) implementation of the CRUNA time-stepping hot path.

The product is libcruna_hip.so (C ABI, include/cruna_hip.h); `cruna` is the
Python host mirror of the reference's interface and `ic` builds the seeded
synthetic workloads.
"""
from . import ic  # noqa: F401
from .cruna import (ADJOINT, DIFF_O2_EXPLICIT, DIFF_O4_EXPLICIT, DIFF_O4_PADE, DIFF_O6_COMPACT, FORWARD, FP32_MIXED, FP64,  # noqa: F401
                    EQ_EULER, EQ_NS, OPEN, PERIODIC, POWER, SUTHERLAND, Closure, Cruna, CrunaError, Ensemble, lib)
