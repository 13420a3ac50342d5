"""Peirce product rule of the oracle vs the reference's table and vs
Gauss-Legendre / Gauss-Jacobi rules from scipy (dg/peirce_quadrature.hpp)."""
import json
import os

import numpy as np
from scipy.special import roots_jacobi, roots_legendre

import oracle_ffi as of

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "peirce_m14.json")


def test_matches_reference_table():
    ref = json.load(open(GOLDEN))["values"]
    got = of.peirce_constants()
    for name in ("cosphi", "sinphi", "b", "r", "c"):
        want = np.array([float(v) for v in ref[name]])
        np.testing.assert_allclose(got[name], want, rtol=0, atol=2e-15, err_msg=name)


def test_matches_scipy_rules():
    got = of.peirce_constants()
    x, w = roots_legendre(30)
    np.testing.assert_allclose(got["cosphi"], x, atol=1e-14)
    np.testing.assert_allclose(got["b"], w, atol=1e-14)
    xj, wj = roots_jacobi(15, 0.0, 0.5)
    np.testing.assert_allclose(got["r"], np.sqrt((xj + 1) / 2), atol=1e-14)
    np.testing.assert_allclose(got["c"], wj / (4 * np.sqrt(2)), atol=1e-14)


def _ball_integral(f):
    """Integral of f over the unit ball with the 60 x 30 x 15 Peirce rule."""
    pc = of.peirce_constants()
    total = 0.0
    for i in range(60):
        th = 2 * np.pi * (i + 1) / 60
        for j in range(30):
            u = np.array([pc["sinphi"][j] * np.cos(th), pc["sinphi"][j] * np.sin(th), pc["cosphi"][j]])
            for k in range(15):
                total += 2 * np.pi * pc["b"][j] * pc["c"][k] / 60 * f(pc["r"][k] * u)
    return total


def test_ball_integrals_closed_form():
    # volume, second moments, a quartic: exact for the product rule
    assert abs(_ball_integral(lambda p: 1.0) - 4 * np.pi / 3) < 1e-12
    assert abs(_ball_integral(lambda p: p[0] ** 2) - 4 * np.pi / 15) < 1e-12
    assert abs(_ball_integral(lambda p: p[2] ** 4) - 4 * np.pi / 35) < 1e-12
    assert abs(_ball_integral(lambda p: p[0] * p[1])) < 1e-12
