"""Seeded case lists for batched sweeps of lid-driven cavities.

The sweep draws, per case, an edge N uniformly from {48, 64, ..., 160} (so the
cases have uneven sizes), a Reynolds number log-uniformly in [100, 3200] and a
lid speed uniformly in [0.02, 0.08].  Every N is a multiple of 16, so 16x16
and 8x8 workgroup footprints fit every case.
"""
from __future__ import annotations

import numpy as np

SWEEP_SEED = 20261203
EDGES = tuple(range(48, 161, 16))


def sweep_cases(n_cases=64, seed=SWEEP_SEED, edges=EDGES, re_range=(100.0, 3200.0), u_range=(0.02, 0.08)):
    """[(nx, ny, nz, Re, u_inf), ...] for a cubic-cavity sweep."""
    rng = np.random.default_rng(seed)
    N = rng.choice(np.asarray(edges), size=n_cases)
    Re = np.exp(rng.uniform(np.log(re_range[0]), np.log(re_range[1]), size=n_cases))
    u = rng.uniform(u_range[0], u_range[1], size=n_cases)
    return [(int(n), int(n), int(n), float(r), float(v)) for n, r, v in zip(N, Re, u)]


def nodes(cases):
    return sum(c[0] * c[1] * c[2] for c in cases)
