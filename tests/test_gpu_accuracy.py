"""Per-field error of FP32 and FP16 scaled-moment storage against the FP64
oracle (lid-driven cavity from rest, Re 1000, lid 0.05, 64^3, 600 steps).

Fields: rho - 1, velocity (u, v, w together), the diagonal and off-diagonal
second moments.  Bounds: FP32 velocity rel-L2 <= 1e-5 (the solver target);
FP16 velocity rel-L2 <= 1e-3 (the storage target; the stored 3u is rounded
stochastically, which keeps this error from growing like a bias: measured
4.9e-4 for D3Q19 and 4.5e-4 for D3Q27; a CPU emulation of the rounding with
uniform random bits in the FP64 oracle gives 5.5e-4 at 600 steps and 6.0e-4
at 1000, against 1.15e-3 and 1.49e-3 with round-to-nearest, which is what
the GPU showed before), and every field's max error below 5e-3 of that
field's largest magnitude.  The table is printed for the record.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

FIELDS = {"rho-1": [0], "velocity": [1, 2, 3], "Pi diag": [4, 7, 9], "Pi offdiag": [5, 6, 8]}


@pytest.mark.parametrize("Q", [19, 27])
def test_per_field_errors(native_lib, Q):
    from cudalbm_amd import Solver

    n, Re, u, steps = 64, 1000.0, 0.05, 600
    ref = O.run(Q, n, n, n, O.omega(u, Re, n), u, O.rest_state(n, n, n), steps)
    out = {}
    for storage in ("fp32", "fp16"):
        with Solver(n, n, n, Re, u, velocity_set=Q, storage=storage) as s:
            s.step(steps)
            got = s.get_moments()
        for name, ks in FIELDS.items():
            a = np.stack([got[k].astype(np.float64) for k in ks])
            b = np.stack([np.asarray(ref[k], np.float64) for k in ks])
            rel = float(np.linalg.norm(a - b) / np.linalg.norm(b))
            mx = float(np.abs(a - b).max())
            out[(storage, name)] = (rel, mx, float(np.abs(b).max()))
    print(f"\nD3Q{Q} {n}^3 Re {Re} {steps} steps: field  storage  rel-L2  max-abs  (max |field|)")
    for (st, name), (rel, mx, scale) in out.items():
        print(f"  {name:10s} {st}  {rel:.3e}  {mx:.3e}  ({scale:.3e})", flush=True)
    assert out[("fp32", "velocity")][0] <= 1e-5
    assert out[("fp16", "velocity")][0] <= 1e-3
    for (st, name), (rel, mx, scale) in out.items():
        assert mx <= 5e-3 * scale, (st, name)
