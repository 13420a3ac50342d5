"""Cubic spline kernel of the oracle (runner.hpp:596-638) vs closed forms."""
import numpy as np
from scipy.integrate import quad

import oracle_ffi as of

H = 0.01


def W(r, h=H):
    k = 8.0 / (np.pi * h ** 3)
    q = r / h
    return np.where(q <= 0.5, k * (6 * q ** 3 - 6 * q ** 2 + 1),
                    np.where(q <= 1, 2 * k * (1 - q) ** 3, 0.0))


def test_known_values():
    k = 8.0 / (np.pi * H ** 3)
    d = np.array([[0, 0, 0], [H, 0, 0], [0, 0.5 * H, 0], [0, 0, 2 * H]], float)
    g, w = of.cubic_kernel(d, H)
    np.testing.assert_allclose(w[0], k, rtol=1e-14)  # W(0) = k
    assert w[1] == 0.0  # W(h) = 0
    np.testing.assert_allclose(w[2], k / 4, rtol=1e-12)  # q = 1/2: both branches give k/4
    assert w[3] == 0.0
    assert np.all(g[0] == 0)  # grad W(0) = 0
    assert np.all(g[3] == 0)


def test_normalised():
    integral, _ = quad(lambda r: 4 * np.pi * r * r * W(r), 0, H, points=[H / 2], epsabs=1e-13)
    assert abs(integral - 1.0) < 1e-10
    # the oracle's W agrees with the closed form on a radial sweep
    r = np.linspace(0, 1.2 * H, 97)
    d = np.stack([r, np.zeros_like(r), np.zeros_like(r)], 1)
    _, w = of.cubic_kernel(d, H)
    np.testing.assert_allclose(w, W(r), rtol=1e-12, atol=1e-6)


def test_gradient_matches_derivative_and_is_odd():
    rng = np.random.default_rng(0)
    d = rng.uniform(-1.1 * H, 1.1 * H, (500, 3))
    g, w = of.cubic_kernel(d, H)
    gm, _ = of.cubic_kernel(-d, H)
    np.testing.assert_allclose(gm, -g, rtol=0, atol=1e-9)
    eps = 1e-7
    for a in range(3):
        e = np.zeros(3)
        e[a] = eps
        _, wp = of.cubic_kernel(d + e, H)
        _, wm = of.cubic_kernel(d - e, H)
        fd = (wp - wm) / (2 * eps)
        np.testing.assert_allclose(g[:, a], fd, rtol=1e-5, atol=1e-3 * np.abs(g).max())


def test_fp32_close_to_fp64():
    rng = np.random.default_rng(1)
    d = rng.uniform(-H, H, (1000, 3))
    g8, w8 = of.cubic_kernel(d, H, 8)
    g4, w4 = of.cubic_kernel(d, H, 4)
    np.testing.assert_allclose(w4, w8, rtol=0, atol=2e-6 * w8.max())
    np.testing.assert_allclose(g4, g8, rtol=0, atol=2e-6 * np.abs(g8).max())
