"""GPU tests of the fused function objects, FP64 reductions and FP16 storage.

Tolerances:
  * k, S (FP32, from FP32 moments): relative 1e-6 against the same formulas
    evaluated in FP64 on the moments the GPU returned;
  * kMean, SMean (FP64 accumulators of FP32 samples): 1e-9 relative against an
    FP64 running mean of the per-step FP32 fields read back after every step;
  * kinetic energy / residual: 1e-9 relative against FP64 host sums of the same
    fields;
  * FP16 scaled-moment storage: velocity rel-L2 <= 1e-3 against the FP64 oracle
    after 500 steps (D3Q19 48^3, Re 400).
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_function_objects_instantaneous(native_lib):
    from cudalbm_amd import Solver
    from cudalbm_amd import native as N

    n = 32
    with Solver(n, n, n, 400.0, 0.05, velocity_set=19) as s:
        s.fo_enable(N.LBM_FO_K | N.LBM_FO_S)
        s.step(40)
        mom = [m.astype(np.float64) for m in s.get_moments()]
        k = s.fo_get(N.LBM_FIELD_K)
        S = s.fo_get(N.LBM_FIELD_S)
        tau = s.tau
    kr = O.k_field(mom)
    Sr = O.S_field(mom, tau)
    assert np.allclose(k, kr, rtol=1e-6, atol=1e-9)
    assert np.allclose(S, Sr, rtol=1e-5, atol=1e-8)
    assert np.abs(k).max() > 1e-3


def test_function_objects_means(native_lib):
    from cudalbm_amd import Solver
    from cudalbm_amd import native as N

    n = 32
    steps = 12
    with Solver(n, n, n, 400.0, 0.05, velocity_set=19) as s:
        s.fo_enable(N.LBM_FO_K | N.LBM_FO_KMEAN | N.LBM_FO_S | N.LBM_FO_SMEAN)
        kmean = np.zeros(n ** 3)
        smean = np.zeros((6, n ** 3))
        for t in range(steps):
            s.step(1)
            k = s.fo_get(N.LBM_FIELD_K)
            S = s.fo_get(N.LBM_FIELD_S)
            kmean = kmean + (k - kmean) / (t + 1)
            smean = smean + (S - smean) / (t + 1)
        km = s.fo_get(N.LBM_FIELD_KMEAN)
        sm = s.fo_get(N.LBM_FIELD_SMEAN)
    assert np.allclose(km, kmean, rtol=1e-9, atol=1e-15)
    assert np.allclose(sm, smean, rtol=1e-9, atol=1e-15)


def test_reductions_fp64(native_lib):
    from cudalbm_amd import Solver

    n = 32
    with Solver(n, n, n, 400.0, 0.05, velocity_set=19) as s:
        s.step(30)
        ke1, r1 = s.reduce()
        m1 = [m.astype(np.float64) for m in s.get_moments()]
        s.step(10)
        ke2, r2 = s.reduce()
        m2 = [m.astype(np.float64) for m in s.get_moments()]
        ke3, r3 = s.reduce()
    u1 = np.stack(m1[1:4]) / 3.0
    u2 = np.stack(m2[1:4]) / 3.0
    ke_ref = 0.5 * np.sum((1.0 + m2[0]) * np.sum(u2 ** 2, axis=0))
    # the previous field is kept in FP32
    r_ref = np.linalg.norm(u2 - u1.astype(np.float32)) / np.linalg.norm(u2)
    assert r1 == 1.0
    assert abs(ke2 - ke_ref) <= 1e-9 * ke_ref
    assert abs(r2 - r_ref) <= 1e-6 * r_ref
    assert r3 <= 1e-7  # nothing moved since the last call (previous field kept in FP32)
    assert ke3 == ke2


def test_fp16_storage_accuracy(native_lib):
    from cudalbm_amd import Solver

    n, Re, u, steps = 48, 400.0, 0.05, 500
    ref = O.run(19, n, n, n, O.omega(u, Re, n), u, O.rest_state(n, n, n), steps)
    with Solver(n, n, n, Re, u, velocity_set=19, storage="fp16") as s:
        s.step(steps)
        got = s.get_moments()
    vr = np.stack(ref[1:4])
    vg = np.stack(got[1:4]).astype(np.float64)
    err = np.linalg.norm(vg - vr) / np.linalg.norm(vr)
    print(f"fp16 velocity rel-L2 {err:.3e}", flush=True)
    assert err <= 1e-3


def test_layout_tiled8_round_trip(native_lib):
    """Upload in the reference 8^3-tiled order, read back in both orders."""
    from cudalbm_amd import Solver
    from cudalbm_amd import native as N

    nx, ny, nz = 32, 16, 24
    rng = np.random.default_rng(5)
    g = [rng.normal(scale=1e-3, size=nx * ny * nz).astype(np.float32) for _ in range(10)]
    # global -> tiled8 on the host
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x, y, z = x.ravel(), y.ravel(), z.ravel()
    t8 = (x % 8) + 8 * ((y % 8) + 8 * ((z % 8) + 8 * ((x // 8) + (nx // 8) * ((y // 8) + (ny // 8) * (z // 8)))))
    tiled = []
    for a in g:
        b = np.empty_like(a)
        b[t8] = a
        tiled.append(b)
    with Solver(nx, ny, nz, 100.0, 0.05, tile=(16, 16, 4)) as s:
        s.set_moments(tiled, layout=N.LBM_LAYOUT_TILED8)
        back_g = s.get_moments(layout=N.LBM_LAYOUT_GLOBAL)
        back_t = s.get_moments(layout=N.LBM_LAYOUT_TILED8)
    for k in range(10):
        assert np.array_equal(back_g[k], g[k])
        assert np.array_equal(back_t[k], tiled[k])


@pytest.mark.parametrize("Q", [19, 27])
def test_fp16_node_words_round_trip(native_lib, Q):
    """FP16 storage keeps the codes of moments 2j and 2j+1 of a node in one
    32-bit word (low / high half).  Moments whose scaled values are FP16 codes
    (code * decode factor) come back exactly, in both layouts, and so does a
    distinct random code per moment and node: a swapped half, a wrong word
    stride or a wrong pair would move codes between moments or nodes."""
    from cudalbm_amd import Solver
    from cudalbm_amd import native as N

    # decode factors of the codec: (rho-1)/cs2, u/cs, Pi_aa and Pi_ab scales
    enc = np.array([3.0, *[1 / np.sqrt(3.0)] * 3, 2 / 3, 1 / 3, 1 / 3, 2 / 3, 1 / 3, 2 / 3])
    dec = (1.0 / enc).astype(np.float32)
    nx, ny, nz = 32, 16, 24
    rng = np.random.default_rng(11)
    codes = [rng.uniform(-0.5, 0.5, nx * ny * nz).astype(np.float16) for _ in range(10)]
    want = [c.astype(np.float32) * dec[k] for k, c in enumerate(codes)]
    with Solver(nx, ny, nz, 100.0, 0.05, velocity_set=Q, storage="fp16", tile=(16, 16, 4)) as s:
        s.set_moments(want)
        back = s.get_moments()
        back_t = s.get_moments(layout=N.LBM_LAYOUT_TILED8)
        s.set_moments(back_t, layout=N.LBM_LAYOUT_TILED8)
        again = s.get_moments()
    for k in range(10):
        assert np.array_equal(back[k], want[k]), k
        assert np.array_equal(again[k], want[k]), k


@pytest.mark.parametrize("Q,storage,plain,means", [
    (19, "fp32", (32, 16), (32, 32)),
    (19, "fp16", (32, 32), (32, 32)),
    (27, "fp32", (32, 16), (16, 16)),
    (27, "fp16", (32, 32), (32, 8)),
])
def test_fo_enable_keeps_the_moments(native_lib, Q, storage, plain, means):
    """Turning the FP64 means on re-tiles an automatically tiled case whose
    fastest footprint changes with them (D3Q19 FP32 32x16 -> 32x32; D3Q27
    32x16 or, FP16, 32x32 with one up-going LDS bank -> 16x16 or 32x8 with two: moments
    permuted on the device, FP16 node words included, face buffers rebuilt) and
    turning them off moves it back; D3Q19 FP16 stays on 32x32.  The moments
    survive both moves bit for bit, a run with both moves equals a run with
    none (results do not depend on the tiling), and the residual's stored
    previous field follows the moves (FP64 sums over another node order: 1e-12
    relative)."""
    from cudalbm_amd import Solver
    from cudalbm_amd import native as N

    n = 32
    with Solver(n, n, n, 400.0, 0.05, velocity_set=Q, storage=storage) as s:
        t0 = s.tiling
        s.step(30)
        s.reduce()
        a = s.get_moments()
        s.fo_enable(N.LBM_FO_KMEAN | N.LBM_FO_SMEAN | N.LBM_FO_MOMENTSMEAN)
        t1 = s.tiling
        b = s.get_moments()
        s.step(5)
        ke1, r1 = s.reduce()
        c = s.get_moments()
        s.fo_enable(0)
        t2 = s.tiling
        c2 = s.get_moments()
        s.step(5)
        ke2, r2 = s.reduce()
        d = s.get_moments()
    with Solver(n, n, n, 400.0, 0.05, velocity_set=Q, storage=storage) as s:
        s.step(30)
        s.reduce()
        s.step(5)
        ke1r, r1r = s.reduce()
        s.step(5)
        ke2r, r2r = s.reduce()
        e = s.get_moments()
    print(f"D3Q{Q} {storage}: tilings {t0} -> {t1} -> {t2}; moments kept across both moves", flush=True)
    assert t0[:2] == plain and t1[:2] == means and t2 == t0
    for x, y in ((ke1, ke1r), (r1, r1r), (ke2, ke2r), (r2, r2r)):
        assert abs(x - y) <= 1e-12 * abs(y), (x, y)
    for k in range(10):
        assert np.array_equal(a[k], b[k]), k
        assert np.array_equal(c[k], c2[k]), k
        assert np.array_equal(d[k], e[k]), k
    assert not np.array_equal(c[1], d[1])


@pytest.mark.parametrize("Q", [19, 27])
def test_function_object_means_against_oracle(native_lib, Q):
    """kMean, SMean and momentsMean accumulated in FP64 inside the GPU step
    against the FP64 oracle's running means of its own trajectory
    (lbmo_run_means), started at absolute step t0 = 100 so that the device step
    clock's 1/(t+1) is exercised away from zero.  Tolerance: rel-L2 per field
    <= 1e-5, the step's parity bar (measured on MI355X: at most 6e-7 for D3Q19
    and 1.5e-6 for D3Q27)."""
    from cudalbm_amd import Solver
    from cudalbm_amd import native as N

    n, Re, u, steps, t0 = 32, 400.0, 0.05, 60, 100
    om = O.omega(u, Re, n)
    mom = O.rest_state(n, n, n)
    ref = O.run_means(Q, n, n, n, om, u, mom, steps, 1.0 / om, t0=t0)
    with Solver(n, n, n, Re, u, velocity_set=Q) as s:
        s.time = t0
        s.fo_enable(N.LBM_FO_KMEAN | N.LBM_FO_SMEAN | N.LBM_FO_MOMENTSMEAN)
        s.step(steps)
        km = s.fo_get(N.LBM_FIELD_KMEAN)
        sm = s.fo_get(N.LBM_FIELD_SMEAN)
        mm = s.fo_get(N.LBM_FIELD_MOMENTSMEAN)

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    errs = {"kMean": rel(km, ref["kMean"])}
    errs.update({f"SMean[{a}]": rel(sm[a], ref["SMean"][a]) for a in range(6)})
    errs.update({f"momentsMean[{a}]": rel(mm[a], ref["momentsMean"][a]) for a in range(10)})
    print(f"D3Q{Q} means vs oracle: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    assert max(errs.values()) <= 1e-5, errs
