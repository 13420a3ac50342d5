"""Pins the oracle's ISPH pressure operator (calculate_isph_diagonal_adv_density
and isph_solve_iteration, runner.hpp:2902-3054) with facts that do not go
through its formula.

One weighted-Jacobi sweep from a chosen pressure field P reports, per
particle, the diagonal (aii), the advected density, the new pressure and the
density error; from them the operator's action (A P)_i = diag_i P_i +
off_i is recovered.  On a perfect lattice at rest, for interior particles,
A is ρ0 (ρ_i + ρ_j)/(ρ_i ρ_j) times the Brookshaw Laplacian
sum_j V_j (x_ij . grad W)/(|x_ij|^2 + 0.01 h^2) (P_i - P_j): it annihilates
constant and linear fields, and for P = s x^2 its continuum limit is
2 s ρ0 / ρ_i (the lattice sum of V x^2 W'(r)/r tends to -1 for a normalised
kernel), to 2 % with the regulariser and the lattice; the mixed field x y
gives nothing.
"""
import numpy as np

import oracle_ffi as of
from alluvion_amd import scenes
from test_oracle_solver import lattice_sum

RB, HB = 2.0 ** -9, 2.0 ** -7  # binary-exact lattice: FP32 positions stay symmetric


def isph_operator(pfun, m=24):
    d = 2 * RB
    L = m * d
    g = np.arange(m) * d - 0.5 * L + 0.5 * d
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    c = int(round(L / HB))
    sc = scenes.Scene(name="isph_lattice", x=x.astype(np.float32), v=np.zeros_like(x, np.float32),
                      particle_radius=RB, kernel_radius=HB, particle_mass=1000.0 / lattice_sum(d, HB),
                      wrap_length=(L, L, L), gravity=(0, 0, 0), viscosity=0.0,
                      boundary_viscosity=0.0, grid_offset=(-c // 2,) * 3, grid_res=(c,) * 3,
                      boundaries=[], pile_gravity=(0, 0, 0), min_density_solve=1,
                      max_density_solve=1)
    orc = of.Oracle(sc, precision=8, kind="i")
    P = pfun(x)
    orc.set("pressure", 2.0 * P)  # the step starts from last_pressure = pressure / 2
    orc.step()
    n, dt = sc.n, sc.dt
    diag = orc.get("aii")[:n]
    b = 1.0 - orc.get("adv_density")[:n]
    p_new = orc.get("pressure")[:n]
    err = orc.get("density_err")[:n]
    assert np.all(p_new > 0)
    off = (err + b - diag * dt * dt * p_new) / (dt * dt)
    inner = np.all(np.abs(x) < 0.5 * L - 2 * HB, axis=1)
    assert inner.sum() >= 512
    return (diag * P + off)[inner], orc.get("density")[:n][inner], diag[inner]


def test_constant_and_linear_fields_are_annihilated():
    ap, _, diag = isph_operator(lambda x: np.full(x.shape[0], 1.0))
    assert np.abs(ap).max() < 1e-9 * np.abs(diag).max()
    ap, _, diag = isph_operator(lambda x: 1.0 + x @ np.array([0.5, -1.0, 2.0]))
    assert np.abs(ap).max() < 1e-9 * np.abs(diag).max()


def test_quadratic_field_gives_the_laplacian():
    s = 3.0
    ap, rho, _ = isph_operator(lambda x: 1.0 + s * x[:, 0] ** 2)
    want = 2.0 * s * 1000.0 / rho
    np.testing.assert_allclose(ap, want, rtol=0.02)
    ap_y, _, _ = isph_operator(lambda x: 1.0 + s * x[:, 1] ** 2)
    np.testing.assert_allclose(ap_y, ap, rtol=1e-9)  # the lattice is isotropic
    ap_xy, _, _ = isph_operator(lambda x: 1.0 + s * x[:, 0] * x[:, 1])
    assert np.abs(ap_xy).max() < 1e-9 * np.abs(ap).max()
