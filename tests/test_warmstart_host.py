"""Host-side pieces of the warm-start tooling (no GPU): the trilinear
upsampling from the network grid and the equilibrium moments of a predicted
field, as the perfect-network bound uses them."""
import numpy as np

from cudalbm_amd import warmstart as W


def test_upsample_exact_for_linear_fields():
    # trilinear interpolation reproduces any function linear in each coordinate
    S, n = 9, 25
    c = np.arange(S) / (S - 1)
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    f = np.stack([1 + 2 * x - y + 0.5 * z, x * y * z, 3 * y, -z], -1).reshape(-1, 4)
    up = W.upsample(f, S, n)
    cn = np.arange(n) / (n - 1)
    zn, yn, xn = np.meshgrid(cn, cn, cn, indexing="ij")
    want = np.stack([1 + 2 * xn - yn + 0.5 * zn, xn * yn * zn, 3 * yn, -zn], -1)
    assert np.allclose(up, want, atol=1e-12)


def test_upsample_then_resample_round_trip():
    # network nodes that fall on solver nodes come back unchanged
    S, n = 8, 15
    rng = np.random.default_rng(3)
    f = rng.normal(size=(S ** 3, 4))
    up = W.upsample(f, S, n)
    back = np.stack([W.resample(up[..., c].reshape(-1), n, S) for c in range(4)], -1)
    assert np.allclose(back, f, atol=1e-12)


def test_equilibrium_moments_walls_and_lid():
    n, U = 6, 0.05
    rng = np.random.default_rng(4)
    f = rng.normal(size=(n, n, n, 4))
    m = [a.reshape(n, n, n) for a in W.equilibrium_moments(f, n, U)]
    # lid row y = n-1 (incl. its edges): u = U along +x, at rest otherwise
    assert np.allclose(m[1][:, n - 1, :], 3 * U) and np.allclose(m[2][:, n - 1, :], 0)
    assert np.allclose(m[4][:, n - 1, :], 4.5 * U * U)
    # the other walls at rest
    for sl in [(slice(None), slice(None), 0), (slice(None), slice(None), n - 1), (0, slice(None), slice(None)),
               (n - 1, slice(None), slice(None)), (slice(None), 0, slice(None))]:
        wall = np.zeros((n, n, n), bool)
        wall[sl] = True
        wall[:, n - 1, :] = False
        for k in range(1, 10):
            assert np.all(m[k][wall] == 0)
    # interior: u = U f0, p = U^2 f3 -> rho - 1 = 3 U^2 f3, Pi_ab from u_a u_b
    i = (slice(1, n - 1), slice(1, n - 2), slice(1, n - 1))
    u, v = U * f[..., 0][i], U * f[..., 1][i]
    assert np.allclose(m[0][i], 3 * U * U * f[..., 3][i], rtol=1e-6)
    assert np.allclose(m[1][i], 3 * u, rtol=1e-6) and np.allclose(m[5][i], 9 * u * v, rtol=1e-5, atol=1e-12)
