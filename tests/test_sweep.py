"""Seeded sweep case lists (host logic only)."""
import numpy as np

from cudalbm_amd.sweep import EDGES, SWEEP_SEED, nodes, sweep_cases


def test_sweep_cases_deterministic_and_in_range():
    a = sweep_cases()
    assert a == sweep_cases(64, SWEEP_SEED)
    assert a != sweep_cases(64, SWEEP_SEED + 1)
    assert len(a) == 64
    for nx, ny, nz, Re, u in a:
        assert nx == ny == nz and nx in EDGES and nx % 16 == 0
        assert 100.0 <= Re <= 3200.0 and 0.02 <= u <= 0.08
    # uneven sizes: several distinct edges, log-uniform Re spans both decades
    assert len({c[0] for c in a}) >= 5
    re = np.array([c[3] for c in a])
    assert re.min() < 200 and re.max() > 1600
    assert nodes(a) == sum(c[0] ** 3 for c in a)
