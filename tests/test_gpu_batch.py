"""GPU tests of the batched multi-case sweep (lbm_batch_*) and the autotuners.

The batch kernel runs the same step body as the single-case kernel, with the
arithmetic fixed by the source (no implicit FMA contraction), so each case
equals a single-case run of that case bit for bit.  Residual/energy
reductions: relative 1e-12 (FP64 sums over different block partitions).  One
case is also checked directly against the FP64 oracle (velocity rel-L2 <=
1e-5).
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = [(32, 32, 32, 100.0, 0.05), (48, 48, 48, 400.0, 0.08), (64, 32, 48, 1000.0, 0.03), (16, 48, 32, 250.0, 0.06)]


def lz_for(nz, target):
    for lz in range(min(nz, target), 1, -1):
        if nz % lz == 0:
            return lz
    return 1


def single(case, steps, tile, Q=19):
    from cudalbm_amd import Solver

    nx, ny, nz, Re, u = case
    with Solver(nx, ny, nz, Re, u, velocity_set=Q, tile=(tile[0], tile[1], lz_for(nz, tile[2]))) as s:
        s.step(steps)
        return s.get_moments()


def assert_same(a, b, what):
    for k in range(10):
        if not np.array_equal(a[k], b[k]):
            d = float(np.abs(a[k].astype(np.float64) - b[k]).max())
            raise AssertionError(f"{what}: moment {k} differs by up to {d:.3e}")


@pytest.mark.parametrize("persistent", [0, 1])
@pytest.mark.parametrize("steps", [37, 130])
def test_batch_matches_single_cases(native_lib, persistent, steps):
    from cudalbm_amd import Batch

    tile = (16, 16, 16)
    print(f"batch persistent={persistent} steps={steps}", flush=True)
    with Batch(CASES, tile=tile, persistent=persistent) as b:
        assert b.tiling == (16, 16, 16, persistent)
        b.step(steps)
        got = [b.get_moments(i) for i in range(len(CASES))]
        assert list(b.steps) == [steps] * len(CASES)
    for i, c in enumerate(CASES):
        assert_same(got[i], single(c, steps, tile), f"case {i} {c}")


@pytest.mark.parametrize("persistent", [0, 1])
def test_batch_matches_single_cases_d3q27(native_lib, persistent):
    """D3Q27 batches (one up-going LDS bank, the persistent form looping over
    items with it) against single-case D3Q27 runs, bit for bit."""
    from cudalbm_amd import Batch

    tile, steps = (16, 16, 16), 41
    print(f"D3Q27 batch persistent={persistent} steps={steps}", flush=True)
    with Batch(CASES, velocity_set=27, tile=tile, persistent=persistent) as b:
        assert b.tiling == (16, 16, 16, persistent)
        b.step(steps)
        got = [b.get_moments(i) for i in range(len(CASES))]
    for i, c in enumerate(CASES):
        assert_same(got[i], single(c, steps, tile, Q=27), f"D3Q27 case {i} {c}")


def test_batch_case_against_oracle(native_lib):
    from cudalbm_amd import Batch

    n, Re, u, steps = 32, 100.0, 0.05, 100
    cases = [(n, n, n, Re, u), (48, 48, 48, 400.0, 0.08)]
    with Batch(cases, tile=(8, 8, 8), persistent=0) as b:
        b.step(steps)
        got = b.get_moments(0)
    ref = O.run(19, n, n, n, O.omega(u, Re, n), u, O.rest_state(n, n, n), steps)
    vr = np.stack(ref[1:4])
    vg = np.stack(got[1:4]).astype(np.float64)
    err = np.linalg.norm(vg - vr) / np.linalg.norm(vr)
    rho_err = np.abs(got[0] - ref[0]).max()
    print(f"batch case vs oracle: velocity rel-L2 {err:.3e}, rho-1 max abs {rho_err:.3e}", flush=True)
    assert err <= 1e-5
    assert rho_err <= 1e-6


def test_batch_set_active_freezes_cases(native_lib):
    """Inactive cases keep fields and step count; odd step counts exercise the
    per-case face parity when a case rejoins."""
    from cudalbm_amd import Batch

    tile = (16, 16, 16)
    with Batch(CASES, tile=tile, persistent=0) as b:
        full = b.work_items
        b.step(11)
        b.set_active([1, 0, 1, 0])
        assert b.work_items < full
        b.step(7)
        assert list(b.steps) == [18, 11, 18, 11]
        mid = [b.get_moments(i) for i in range(4)]
        b.set_active([1, 1, 0, 1])
        b.step(5)
        assert list(b.steps) == [23, 16, 18, 16]
        end = [b.get_moments(i) for i in range(4)]
    assert_same(mid[1], single(CASES[1], 11, tile), "frozen case 1")
    assert_same(mid[0], single(CASES[0], 18, tile), "case 0 at 18")
    for i, t in enumerate([23, 16, 18, 16]):
        assert_same(end[i], single(CASES[i], t, tile), f"case {i} at {t}")


def test_batch_reduce_matches_single(native_lib):
    from cudalbm_amd import Batch, Solver

    tile = (16, 16, 16)
    with Batch(CASES, tile=tile, persistent=0) as b:
        b.step(20)
        ke1, r1 = b.reduce()
        b.step(10)
        ke2, r2 = b.reduce()
    assert np.all(r1 == 1.0)
    for i, (nx, ny, nz, Re, u) in enumerate(CASES):
        with Solver(nx, ny, nz, Re, u, tile=(16, 16, lz_for(nz, 16))) as s:
            s.step(20)
            k1, q1 = s.reduce()
            s.step(10)
            k2, q2 = s.reduce()
        # equal fields; the FP64 sums may run over another block partition
        assert abs(ke1[i] - k1) <= 1e-12 * k1
        assert abs(ke2[i] - k2) <= 1e-12 * k2
        assert abs(r2[i] - q2) <= 1e-12 * q2


def test_batch_run_stops_each_case(native_lib):
    from cudalbm_amd import Batch

    tol, check, max_steps = 0.1, 40, 400  # at step 400 the 40-step residuals are 0.05-0.15
    with Batch(CASES, tile=(16, 16, 16), persistent=0) as b:
        steps, res, ms = b.run(tol, check, max_steps)
        assert b.work_items == 0  # nothing left active
        assert list(b.steps) == list(steps)
    print("run:", list(steps), [f"{r:.3g}" for r in res], f"{ms:.1f} ms", flush=True)
    assert ms > 0
    for s, r in zip(steps, res):
        assert 0 < s <= max_steps
        assert r < tol or s == max_steps
        # a case stops at the first check below tol
        assert s % check == 0 or s == max_steps
    assert any(r < tol for r in res)


def test_batch_autotune_matches_fixed(native_lib):
    """Tuned tilings change only the work decomposition: fields agree to FP32 rounding."""
    from cudalbm_amd import Batch

    cases = CASES[:3]
    with Batch(cases, tile=(0, 0, 0), persistent=-1) as b:
        rep = b.tuning_report
        tx, ty, tz, pers = b.tiling
        b.step(24)
        got = [b.get_moments(i) for i in range(len(cases))]
    print(rep, flush=True)
    lines = [l for l in rep.splitlines() if "ms/step" in l]
    assert len(lines) >= 4
    assert (tx, ty) in [(16, 16), (8, 8)]  # the footprints that fit every case
    for i, c in enumerate(cases):
        assert_same(got[i], single(c, 24, (16, 16, 16)), f"case {i}")


def test_single_case_autotune(native_lib):
    from cudalbm_amd import Solver, autotune

    (tx, ty, tz), ms = autotune(64, 64, 64, 400.0, 0.05, steps=10)
    assert (tx, ty) in [(32, 32), (64, 16), (32, 16), (16, 16), (32, 8), (8, 8)] and 64 % tz == 0 and ms > 0
    with Solver(64, 64, 64, 400.0, 0.05, tile=(tx, ty, tz)) as s:
        assert s.tiling == (tx, ty, tz)


def test_batch_errors(native_lib):
    from cudalbm_amd import Batch, LBMError

    with pytest.raises(LBMError, match="-49"):
        Batch([(40, 40, 40, 100.0, 0.05)], tile=(16, 16, 16), persistent=0)
    with pytest.raises(LBMError, match="-1"):
        Batch([(32, 32, 32, -1.0, 0.05)], tile=(16, 16, 16), persistent=0)
    with Batch([(32, 32, 32, 100.0, 0.05)], tile=(16, 16, 16), persistent=0) as b:
        with pytest.raises(LBMError, match="out of range"):
            b.get_moments(3)
