"""Synthetic ensemble workload: many periodic boxes of uneven sizes.

`ensemble_members` draws the members of a batched study: box sizes N from
{48, 64, 96, 128} (N^3 points each) and rms Mach numbers from U[0.05, 0.3]
with one seeded generator; member i gets the seeded solenoidal initial
field of `ic.periodic_box_state` with seed `seed + i`, Re 1600 with the
power-law viscosity mu0 (T/T0)**0.76 (T0 the initial temperature), and dt
from CFL 0.5.  The first half of the members are forward RK4 runs; the
second half are adjoint runs backward in time over a forward trajectory
kept as checkpoints every `subset_steps` steps (the reference's time.subsets
cache, src/main/computation_direct.f90:55-133 and
src/main/computation_adjoint.f90:45-135), with random final adjoint states.
"""
import numpy as np

from . import ic
from .cruna import POWER


def ensemble_members(nmembers=32, seed=0, sizes=(48, 64, 96, 128), mach_range=(0.05, 0.3)):
    """(specs, states, adjoint): Cruna keyword dicts (n, dx, ...), initial
    states and the set of adjoint member indices."""
    rng = np.random.default_rng(seed)
    ns = rng.choice(list(sizes), nmembers)
    machs = rng.uniform(mach_range[0], mach_range[1], nmembers)
    specs, states = [], []
    for i in range(nmembers):
        n = int(ns[i])
        q = ic.periodic_box_state(n, seed + i, float(machs[i]))
        h = ic.periodic_dx(n)
        specs.append(dict(n=(n,) * 3, dx=(h,) * 3, mu0=1.0 / 1600.0, T0=1.0 / (1.4 * machs[i] ** 2),
                          dt=ic.cfl_dt(q, h), visc_law=POWER, visc_exp=0.76, afilter_freq=(1, 1, 1)))
        states.append(q)
    adjoint = set(range(nmembers // 2, nmembers))
    return specs, states, adjoint


def final_adjoint_state(shape, seed, i):
    """The final-time adjoint state of member i (seeded normal field)."""
    return np.asfortranarray(np.random.default_rng(seed + 100 + i).standard_normal(shape))
