"""The U-Net oracle's hand-written backward against torch autograd (CPU only).

Pins oracle/unet_ref.py: without bfloat16 rounding its manual gradient must
equal autograd's of the same forward to 1e-10 relative (float64)."""
import numpy as np
import torch

from oracle import unet_ref as R


def rel(a, b):
    return float(torch.linalg.norm((a - b).reshape(-1)) / max(float(torch.linalg.norm(b.reshape(-1))), 1e-300))


def test_manual_backward_matches_autograd():
    rng = np.random.default_rng(3)
    B, S = 2, 8
    params = R.unpack(rng.normal(scale=0.2, size=R.num_params()))
    x = R.pad_input(rng.normal(size=(B, S ** 3, 3)), B, S)
    t = torch.from_numpy(rng.normal(size=(B, 4, S, S, S)))
    l1, g1 = R.loss_and_grad(params, x, t, bf16=False)
    l2, g2 = R.autograd_loss_and_grad(params, x, t)
    assert abs(l1 - l2) <= 1e-12 * abs(l2)
    for i, ((dw1, db1), (dw2, db2)) in enumerate(zip(g1, g2)):
        assert rel(dw1, dw2) <= 1e-10, i
        assert rel(db1, db2) <= 1e-10, i


def test_pack_unpack_round_trip():
    rng = np.random.default_rng(4)
    flat = rng.normal(size=R.num_params())
    params = R.unpack(flat)
    assert np.array_equal(R.pack_grads(params), flat)
    # weight [co][t][ci] -> torch [co][ci][kz][ky][kx] with t = 9 kz + 3 ky + kx
    w0 = flat[:16 * 27 * 8].reshape(16, 27, 8)
    assert params[0][0][5, 3, 2, 1, 0] == w0[5, 2 * 9 + 1 * 3 + 0, 3]


def test_bf16_rounding_is_nearest_even():
    # ulp(1) = 2^-7 in bfloat16: below half an ulp rounds down, above rounds up,
    # ties go to the even mantissa
    x = torch.tensor([1.0 + 2.0 ** -9, 1.0 + 3 * 2.0 ** -9, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8],
                     dtype=torch.float64)
    assert R.bf(x).tolist() == [1.0, 1.0 + 2.0 ** -7, 1.0, 1.0 + 2.0 ** -6]
