"""GPU tests of the U-Net surrogate: BF16 MFMA forward and gradients against
the CPU oracle (oracle/unet_ref.py, float64 with BF16 rounding at the same
points as the GPU), training, and the warm-start conversion.

Tolerances: the GPU accumulates in FP32 in a different order than the float64
oracle, so a few values land on the other side of a BF16 rounding step and
propagate; forward output rel-L2 <= 1e-2, per-layer weight and bias gradient
rel-L2 <= 3e-2, loss relative 1e-3.  A wrong index, tap or transpose gives
O(1) errors.  Warm-start moments: 1e-6 absolute against a host recomputation
of the same interpolation and equilibrium conversion.
"""
import numpy as np
import pytest
import torch

from oracle import unet_ref as R

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.ravel(a).astype(np.float64), np.ravel(b).astype(np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def case_inputs(S, cases):
    from cudalbm_amd.native import encode_inputs

    return np.stack([encode_inputs(S, Re, N) for Re, N in cases])


def smooth_targets(S, cases):
    """Smooth fields standing in for cavity solutions: [n][S^3][4]."""
    z, y, x = np.meshgrid(*(np.linspace(0, 1, S),) * 3, indexing="ij")
    out = []
    for Re, N in cases:
        a = np.log(Re / 800.0)
        u = np.sin(np.pi * x) * y ** 3 * (1 + 0.2 * a)
        v = -0.5 * np.sin(2 * np.pi * x) * np.sin(np.pi * y) * (1 - 0.1 * a)
        w = 0.1 * np.sin(np.pi * z) * np.sin(np.pi * x) * y
        p = 0.3 * (x - 0.5) * y ** 2
        out.append(np.stack([u, v, w, p], -1).reshape(-1, 4))
    return np.stack(out).astype(np.float32)


def make_net(S, seed):
    from cudalbm_amd import Surrogate

    net = Surrogate(S=S, seed=seed)
    # non-zero biases so the bias path is exercised
    p = net.get_params()
    rng = np.random.default_rng(seed)
    nw = sum(co * t * ci for ci, co, t, _ in R.CONVS)
    p[nw:] = rng.normal(scale=0.05, size=p.size - nw)
    # the head's unused outputs stay zero
    p[-4:] = 0.0
    net.set_params(p)
    return net, p


def test_forward_matches_oracle(native_lib):
    S, cases = 16, [(400.0, 96), (2500.0, 160)]
    x = case_inputs(S, cases)
    x[..., 1:] += np.random.default_rng(1).normal(scale=0.3, size=x[..., 1:].shape).astype(np.float32)
    net, p = make_net(S, 11)
    with net:
        got = net.predict(x)
    out, _ = R.forward(R.unpack(p), R.pad_input(x, len(cases), S), bf16=True)
    ref = R.from_ncdhw(out[:, :4]).reshape(len(cases), S ** 3, 4)
    err = rel(got, ref)
    print(f"forward rel-L2 {err:.3e}, output scale {np.abs(ref).max():.3e}", flush=True)
    assert np.abs(ref).max() > 1e-3
    assert err <= 1e-2


def test_gradient_matches_oracle(native_lib):
    S, cases = 16, [(150.0, 48), (1200.0, 128)]
    x = case_inputs(S, cases)
    t = smooth_targets(S, cases)
    net, p = make_net(S, 12)
    with net:
        loss, g = net.gradient(x, t)
    params = R.unpack(p)
    tt = R.to_ncdhw(t, len(cases), S, 4)
    lref, gref = R.loss_and_grad(params, R.pad_input(x, len(cases), S), tt, bf16=True)
    print(f"loss gpu {loss:.6e} oracle {lref:.6e}", flush=True)
    assert abs(loss - lref) <= 1e-3 * lref
    gflat = R.pack_grads(gref)
    off_w, off_b = 0, sum(co * t_ * ci for ci, co, t_, _ in R.CONVS)
    worst = 0.0
    for i, (ci, co, taps, _) in enumerate(R.CONVS):
        nw = co * taps * ci
        ew = rel(g[off_w:off_w + nw], gflat[off_w:off_w + nw])
        eb = rel(g[off_b:off_b + co], gflat[off_b:off_b + co])
        print(f"conv {i:2d} ({ci:3d}->{co:3d}): dW rel-L2 {ew:.2e}  db rel-L2 {eb:.2e}", flush=True)
        worst = max(worst, ew, eb)
        off_w += nw
        off_b += co
    assert worst <= 3e-2


def test_training_reduces_loss(native_lib):
    from cudalbm_amd import Surrogate

    S = 16
    rng = np.random.default_rng(5)
    cases = [(float(np.exp(rng.uniform(np.log(100), np.log(3200)))), int(rng.choice([48, 96, 160])))
             for _ in range(16)]
    x = case_inputs(S, cases)
    t = smooth_targets(S, cases)
    with Surrogate(S=S, seed=3) as net:
        net.load_data(x, t)
        losses = net.train(epochs=60, batch=4, lr=2e-3, seed=1)
        pred = net.predict(x[:2])
    print("loss per epoch:", [f"{v:.2e}" for v in losses[::10]], f"final {losses[-1]:.2e}", flush=True)
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.1 * losses[0]
    assert rel(pred, t[:2]) < 0.5


def trilinear_moments(f, S, N, U):
    """Host restatement of lbm_sur_warm_start's interpolation and conversion."""
    f = f.reshape(S, S, S, 4).astype(np.float64)
    sc = np.arange(N) * (S - 1) / (N - 1)
    i0 = np.minimum(S - 2, np.floor(sc).astype(int))
    fr = sc - i0
    out = np.zeros((N, N, N, 4))
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                wz = (fr if dz else 1 - fr)[:, None, None]
                wy = (fr if dy else 1 - fr)[None, :, None]
                wx = (fr if dx else 1 - fr)[None, None, :]
                out += (wz * wy * wx)[..., None] * f[np.ix_(i0 + dz, i0 + dy, i0 + dx)]
    u, v, w = U * out[..., 0], U * out[..., 1], U * out[..., 2]
    rm1 = 3 * U * U * out[..., 3]
    wall = np.zeros((N, N, N), bool)
    wall[:, :, 0] = wall[:, :, -1] = wall[0] = wall[-1] = wall[:, 0] = True
    for a in (u, v, w):
        a[wall] = 0
    u[:, -1] = U
    v[:, -1] = 0
    w[:, -1] = 0
    m = [rm1, 3 * u, 3 * v, 3 * w, 4.5 * u * u, 9 * u * v, 9 * u * w, 4.5 * v * v, 9 * v * w, 4.5 * w * w]
    return [a.reshape(-1) for a in m]


def test_warm_start_moments(native_lib):
    from cudalbm_amd import Solver
    from cudalbm_amd.native import encode_inputs

    S, N, Re, U = 16, 24, 700.0, 0.06
    net, _ = make_net(S, 13)
    with net, Solver(N, N, N, Re, U, tile=(8, 8, 0)) as s:
        f = net.predict(encode_inputs(S, Re, N)[None])[0]
        net.warm_start(s, Re, U)
        got = s.get_moments()
    ref = trilinear_moments(f, S, N, U)
    for k in range(10):
        assert np.abs(got[k] - ref[k]).max() <= 1e-6, k
    assert np.abs(got[1]).max() > 0.1  # the lid row carries 3U
