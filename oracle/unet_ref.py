"""CPU oracle of the 3D U-Net surrogate (cudalbm_amd/csrc/surrogate.hip).

TEST INFRASTRUCTURE ONLY: the checker of the GPU network's forward pass and
gradients.  Nothing in the product loads it.

No reference counterpart exists (the reference has no learned component), so
this restatement is pinned by library routines instead: the convolutions are
torch's CPU float64 `conv3d` and its `grad.conv3d_input` / `conv3d_weight`, and
the hand-written backward here is checked against torch autograd of the same
forward (tests/test_unet_oracle.py).  With `bf16=True` it rounds to bfloat16
at exactly the points where the GPU stores BF16: the weights, the input, every
ReLU output, the pooled and concatenated tensors, and the gradient at every
pre-activation; everything else is float64.

Tensors here are torch NCDHW (D = z, H = y, W = x); the GPU's channels-last
[b][z][y][x][c] arrays are converted at the boundary.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

# (cin, cout, taps, relu) of the 15 convolutions, in parameter order
CONVS = [(8, 16, 27, 1), (16, 16, 27, 1), (16, 32, 27, 1), (32, 32, 27, 1), (32, 64, 27, 1), (64, 64, 27, 1),
         (64, 128, 27, 1), (128, 128, 27, 1), (192, 64, 27, 1), (64, 64, 27, 1), (96, 32, 27, 1), (32, 32, 27, 1),
         (48, 16, 27, 1), (16, 16, 27, 1), (16, 8, 1, 0)]
IN_USED, IN_PAD, OUT_USED = 3, 8, 4


def num_params():
    return sum(co * t * ci + co for ci, co, t, _ in CONVS)


def bf(x, on=True):
    return x.to(torch.bfloat16).to(torch.float64) if on else x


def unpack(flat):
    """Flat FP32 parameters (all weights [co][t][ci], then all biases) -> [(W, b)] in torch shapes."""
    flat = np.asarray(flat, dtype=np.float64)
    out, off = [], 0
    ws = []
    for ci, co, t, _ in CONVS:
        n = co * t * ci
        w = flat[off:off + n].reshape(co, t, ci)
        k = 3 if t == 27 else 1
        ws.append(torch.from_numpy(np.ascontiguousarray(w.transpose(0, 2, 1).reshape(co, ci, k, k, k))))
        off += n
    for i, (ci, co, t, _) in enumerate(CONVS):
        out.append((ws[i], torch.from_numpy(flat[off:off + co].copy())))
        off += co
    return out


def pack_grads(grads):
    """[(dW, db)] in torch shapes -> flat vector in the GPU parameter order."""
    parts = []
    for (dw, _), (ci, co, t, _) in zip(grads, CONVS):
        parts.append(dw.reshape(co, ci, t).permute(0, 2, 1).reshape(-1).numpy())
    for (_, db) in grads:
        parts.append(db.numpy())
    return np.concatenate(parts)


def to_ncdhw(a, B, S, C):
    return torch.from_numpy(np.asarray(a, dtype=np.float64).reshape(B, S, S, S, C)).permute(0, 4, 1, 2, 3).contiguous()


def from_ncdhw(t):
    return t.permute(0, 2, 3, 4, 1).contiguous().numpy()


def _conv(x, w, b, relu):
    pad = 1 if w.shape[-1] == 3 else 0
    y = F.conv3d(x, w, b, padding=pad)
    return torch.relu(y) if relu else y


def forward(params, x_in, bf16=True):
    """x_in: torch (B, 8, S, S, S) padded input.  Returns (output (B, 8, S, S, S), cache)."""
    Wb = [(bf(w, bf16), b) for w, b in params]
    c = {"x0": bf(x_in, bf16)}

    def conv(i, x):
        y = _conv(x, Wb[i][0], Wb[i][1], CONVS[i][3])
        return bf(y, bf16) if CONVS[i][3] else y

    def pool(x):
        return bf(F.avg_pool3d(x, 2), bf16)

    def upcat(low, skip):
        up = low.repeat_interleave(2, 2).repeat_interleave(2, 3).repeat_interleave(2, 4)
        return torch.cat([up, skip], 1)

    c["a1"] = conv(0, c["x0"]); c["s1"] = conv(1, c["a1"]); c["p1"] = pool(c["s1"])
    c["a2"] = conv(2, c["p1"]); c["s2"] = conv(3, c["a2"]); c["p2"] = pool(c["s2"])
    c["a3"] = conv(4, c["p2"]); c["s3"] = conv(5, c["a3"]); c["p3"] = pool(c["s3"])
    c["b1"] = conv(6, c["p3"]); c["b2"] = conv(7, c["b1"])
    c["u3"] = upcat(c["b2"], c["s3"]); c["d3a"] = conv(8, c["u3"]); c["d3b"] = conv(9, c["d3a"])
    c["u2"] = upcat(c["d3b"], c["s2"]); c["d2a"] = conv(10, c["u2"]); c["d2b"] = conv(11, c["d2a"])
    c["u1"] = upcat(c["d2b"], c["s1"]); c["d1a"] = conv(12, c["u1"]); c["d1b"] = conv(13, c["d1a"])
    c["out"] = conv(14, c["d1b"])
    c["W"] = Wb
    return c["out"], c


def loss_and_grad(params, x_in, target, bf16=True):
    """MSE over output channels 0-3 and its gradient, by hand (mirrors the GPU order of rounding).

    x_in (B, 8, S, S, S), target (B, 4, S, S, S).  Returns (loss, [(dW, db)])."""
    out, c = forward(params, x_in, bf16)
    W = c["W"]
    n = target.numel()
    d = out[:, :OUT_USED] - target
    loss = float((d * d).sum() / n)
    g_out = torch.zeros_like(out)
    g_out[:, :OUT_USED] = 2.0 * d / n
    grads = [None] * len(CONVS)

    def conv_back(i, x, g):
        """g: gradient at the pre-activation (already rounded).  Returns dX."""
        w = W[i][0]
        pad = 1 if w.shape[-1] == 3 else 0
        dw = torch.nn.grad.conv3d_weight(x, w.shape, g, padding=pad)
        db = g.sum(dim=(0, 2, 3, 4))
        grads[i] = (dw, db)
        return torch.nn.grad.conv3d_input(x.shape, w, g, padding=pad)

    def pre(dy, y):
        return bf(dy * (y > 0), bf16)

    def pool_back(dp):
        return dp.repeat_interleave(2, 2).repeat_interleave(2, 3).repeat_interleave(2, 4) / 8.0

    def upcat_back(du, ca):
        dlow = F.avg_pool3d(du[:, :ca], 2) * 8.0
        return dlow, du[:, ca:]

    g = bf(g_out, bf16)
    dd1b = conv_back(14, c["d1b"], g)
    dd1a = conv_back(13, c["d1a"], pre(dd1b, c["d1b"]))
    du1 = conv_back(12, c["u1"], pre(dd1a, c["d1a"]))
    dd2b, ds1 = upcat_back(du1, 32)
    dd2a = conv_back(11, c["d2a"], pre(dd2b, c["d2b"]))
    du2 = conv_back(10, c["u2"], pre(dd2a, c["d2a"]))
    dd3b, ds2 = upcat_back(du2, 64)
    dd3a = conv_back(9, c["d3a"], pre(dd3b, c["d3b"]))
    du3 = conv_back(8, c["u3"], pre(dd3a, c["d3a"]))
    db2, ds3 = upcat_back(du3, 128)
    db1 = conv_back(7, c["b1"], pre(db2, c["b2"]))
    dp3 = conv_back(6, c["p3"], pre(db1, c["b1"]))
    ds3 = ds3 + pool_back(dp3)
    da3 = conv_back(5, c["a3"], pre(ds3, c["s3"]))
    dp2 = conv_back(4, c["p2"], pre(da3, c["a3"]))
    ds2 = ds2 + pool_back(dp2)
    da2 = conv_back(3, c["a2"], pre(ds2, c["s2"]))
    dp1 = conv_back(2, c["p1"], pre(da2, c["a2"]))
    ds1 = ds1 + pool_back(dp1)
    da1 = conv_back(1, c["a1"], pre(ds1, c["s1"]))
    conv_back(0, c["x0"], pre(da1, c["a1"]))
    return loss, grads


def autograd_loss_and_grad(params, x_in, target):
    """Same loss by torch autograd (no rounding): pins the hand-written backward."""
    P = [(w.clone().requires_grad_(True), b.clone().requires_grad_(True)) for w, b in params]
    out, _ = forward(P, x_in, bf16=False)
    n = target.numel()
    loss = ((out[:, :OUT_USED] - target) ** 2).sum() / n
    loss.backward()
    return float(loss), [(w.grad, b.grad) for w, b in P]


def pad_input(x3, B, S):
    """GPU input layout [B][S^3][3] -> torch (B, 8, S, S, S) with zero padding channels."""
    x = to_ncdhw(x3, B, S, IN_USED)
    return torch.cat([x, torch.zeros(B, IN_PAD - IN_USED, S, S, S, dtype=torch.float64)], 1)
