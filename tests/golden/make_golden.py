"""Writes tests/golden/cores.npz: outputs of the reference's own Fortran
derivative (6th-order compact, 4th-order explicit, 4th-order Pade, 2nd-order
explicit) and filter cores
(oracle/_ref, compiled from the reference sources)
on seeded inputs.  Run in the build container where the reference sources
exist:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from tests import refproc  # noqa: E402


def main():
    out = {}
    cases = [("diff", s, n, 0.21) for s in (0, 1) for n in (8, 16, 48)]
    cases += [("filter", s, n, 0.35) for s in (0, 1) for n in (8, 16, 48)]
    cases += [("diffo4", s, n, 0.13) for s in (0, 1) for n in (8, 16, 48)]
    cases += [("diffpade", s, n, 0.19) for s in (0, 1) for n in (8, 16, 48)]
    cases += [("diffo2", 1, n, 0.23) for n in (8, 16, 48)]  # the o2 core is open only
    for kind, scheme, n, arg in cases:
        q = np.asfortranarray(np.random.default_rng(1000 + n + 7 * scheme).standard_normal((n, 4)))
        key = f"{kind}_{scheme}_{n}"
        out[key + "_in"] = q
        out[key + "_arg"] = np.float64(arg)
        ref_kind = {"diffo4": "diff_o4", "diffpade": "diff_pade", "diffo2": "diff_o2"}.get(kind, kind)
        out[key + "_out"] = refproc.run("core", ref_kind, scheme, q, arg)
    np.savez_compressed(os.path.join(os.path.dirname(__file__), "cores.npz"), **out)
    print("wrote", len(cases), "cases")


if __name__ == "__main__":
    main()
