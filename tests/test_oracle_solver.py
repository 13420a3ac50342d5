"""Invariants and closed forms that pin the oracle's DFSPH / IISPH step and
rigid-body integration (no reference test covers them):

* a perfect periodic lattice at rest: density equals the closed lattice sum,
  nothing moves, both solves stop at their minimum iteration counts;
* total momentum of a periodic fluid with no gravity and no viscosity is
  conserved by the pressure solves (pairwise antisymmetric updates);
* an isolated particle falls freely: v += g dt, then x += dt v;
* the min / max iteration semantics of the convergence loops;
* a rigid body far from the fluid: the reference's integrate_kinematics
  closed form (pile.hpp:794-815, calculate_dq runner.hpp:240-250).
"""
import numpy as np
import pytest

import oracle_ffi as of
from alluvion_amd import scenes

R, H = 0.0025, 0.01


def cubic_w(q, h):
    k = 8.0 / (np.pi * h ** 3)
    return np.where(q <= 0.5, 6 * q ** 3 - 6 * q ** 2 + 1, np.where(q <= 1, 2 * (1 - q) ** 3, 0)) * k


def lattice_sum(d, h):
    """W(0) + sum of W over the other points of a cubic lattice of spacing d."""
    m = int(np.ceil(h / d)) + 1
    g = np.arange(-m, m + 1) * d
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2).ravel()
    return float(cubic_w(r / h, h).sum())


def periodic_lattice(m=20, jitter=0.0, vamp=0.0, seed=1, visc=0.0, mass=None, gravity=(0, 0, 0)):
    d = 2 * R
    L = m * d  # m = 20 -> 0.1 m = 10 cells of h
    idx = np.arange(m) * d - 0.5 * L + 0.5 * d
    X, Y, Z = np.meshgrid(idx, idx, idx, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    rng = np.random.default_rng(seed)
    x = x + rng.uniform(-jitter, jitter, x.shape)
    v = rng.uniform(-vamp, vamp, x.shape)
    c = int(round(L / H))
    sc = scenes.Scene(name="periodic_lattice", x=x.astype(np.float32), v=v.astype(np.float32),
                      wrap_length=(L, L, L), gravity=gravity, viscosity=visc,
                      boundary_viscosity=visc, grid_offset=(-c // 2,) * 3, grid_res=(c,) * 3,
                      boundaries=[], pile_gravity=gravity)
    if mass is not None:
        sc.particle_mass = mass
    return sc, x, v


def oracle64(sc, kind="df", xv=None):
    orc = of.Oracle(sc, precision=8, kind=kind)
    if xv is not None:
        orc.set("x", xv[0])
        orc.set("v", xv[1])
    return orc


@pytest.mark.parametrize("kind", ["df", "ii", "i"])
def test_lattice_at_rest(kind):
    S = lattice_sum(2 * R, H)
    mass = 0.99 * 1000.0 / S  # 1 % under rest density
    sc, x, v = periodic_lattice(mass=mass)
    orc = oracle64(sc, kind, (x, v))
    for _ in range(3):
        orc.step()
    rep = orc.report()
    rho = orc.get("density")[: sc.n]
    # closed lattice sum: rho = m (W(0) + sum_j W)
    np.testing.assert_allclose(rho, mass * S, rtol=1e-12)
    assert np.all(orc.get("v")[: sc.n] == 0.0)
    np.testing.assert_array_equal(orc.get("x")[: sc.n], x)
    assert rep["num_density_solve"] == sc.min_density_solve
    if kind == "df":
        assert rep["num_divergence_solve"] == sc.min_divergence_solve
    else:
        assert np.all(orc.get("pressure")[: sc.n] == 0.0)


@pytest.mark.parametrize("kind", ["df", "ii", "i"])
def test_momentum_conserved_without_gravity_and_viscosity(kind):
    sc, x, v = periodic_lattice(jitter=0.3 * R, vamp=0.05, seed=3)
    orc = oracle64(sc, kind, (x, v))
    p0 = orc.get("v")[: sc.n].sum(0)
    v0 = orc.get("v")[: sc.n].copy()
    for _ in range(3):
        orc.step()
    v1 = orc.get("v")[: sc.n]
    assert np.abs(v1 - v0).max() > 1e-4  # the pressure solves did act
    scale = np.abs(v1).sum()
    np.testing.assert_allclose(v1.sum(0), p0, rtol=0, atol=1e-12 * scale)


def test_isolated_particles_fall_freely():
    x = np.array([[0.0, 0.0, 0.0], [0.05, 0.02, -0.03], [-0.04, 0.01, 0.05]])
    v = np.array([[0.1, 0.2, -0.3], [0.0, 0.0, 0.0], [-0.5, 0.1, 0.05]])
    sc = scenes.Scene(name="free", x=x.astype(np.float32), v=v.astype(np.float32),
                      grid_offset=(-10, -10, -10), grid_res=(20, 20, 20), boundaries=[])
    orc = oracle64(sc, "df", (x, v))
    g = np.array(sc.gravity)
    for _ in range(5):
        orc.step()
        v = v + sc.dt * g
        x = x + sc.dt * v
    np.testing.assert_allclose(orc.get("v")[:3], v, rtol=0, atol=1e-15)
    np.testing.assert_allclose(orc.get("x")[:3], x, rtol=0, atol=1e-15)


def test_iteration_count_semantics():
    sc = scenes.dam_break_small()
    sc.boundaries[0].resolution = (8, 8, 8)
    # min counts force iterations past convergence
    sc.min_density_solve, sc.min_divergence_solve = 5, 4
    orc = of.Oracle(sc, precision=8)
    orc.build_volume_map(0)
    orc.step()
    rep = orc.report()
    assert rep["num_density_solve"] == 5 and rep["num_divergence_solve"] == 4
    # max counts stop a solve that cannot reach the tolerance
    sc = scenes.dam_break_small()
    sc.boundaries[0].resolution = (8, 8, 8)
    sc.density_error_tolerance = 1e-12
    sc.max_density_solve = 3
    orc = of.Oracle(sc, precision=8)
    orc.build_volume_map(0)
    for _ in range(3):
        orc.step()
    rep = orc.report()
    assert rep["num_density_solve"] == 3
    assert rep["mean_density_error"] > 1e-12


def quat_dq(w, q):
    """0.5 (0, w) x q, Hamilton product, q = (x, y, z, w)."""
    return 0.5 * np.array([w[0] * q[3] + w[1] * q[2] - w[2] * q[1],
                           -w[0] * q[2] + w[1] * q[3] + w[2] * q[0],
                           w[0] * q[1] - w[1] * q[0] + w[2] * q[3],
                           -w[0] * q[0] - w[1] * q[1] - w[2] * q[2]])


def test_rigid_body_away_from_fluid_follows_the_kinematics_closed_form():
    body = scenes.BoundarySpec("sphere", (0.05,), 1.0, (8, 8, 8), mass=0.3,
                               inertia=(3e-4, 3e-4, 3e-4), x=(0.3, 0.2, -0.1))
    x = np.array([[-0.3, -0.3, -0.3]])
    sc = scenes.Scene(name="body", x=x.astype(np.float32), v=np.zeros((1, 3), np.float32),
                      grid_offset=(-51, -51, -51), grid_res=(102, 102, 102), boundaries=[body])
    orc = of.Oracle(sc, precision=8)
    orc.build_volume_map(0)
    X0, V0 = np.array([0.3, 0.2, -0.1]), np.array([0.1, -0.2, 0.3])
    Q0 = np.array([0.1, 0.2, 0.3, 0.9]) / np.linalg.norm([0.1, 0.2, 0.3, 0.9])
    W0 = np.array([0.0, 0.0, 2.0])
    orc.set_rigid(0, X0, V0, Q0, W0)
    orc.step()
    dt, g = sc.dt, np.array(sc.pile_gravity)
    rb = orc.rigid(0)
    assert np.all(rb["force"] == 0) and np.all(rb["torque"] == 0)
    dx = (g * dt + V0) * dt
    np.testing.assert_allclose(rb["x"], X0 + dx, rtol=0, atol=1e-15)
    np.testing.assert_allclose(rb["v"], dx / dt, rtol=0, atol=1e-13)
    q = Q0 + dt * quat_dq(W0, Q0)
    np.testing.assert_allclose(rb["q"], q / np.linalg.norm(q), rtol=0, atol=1e-15)
    np.testing.assert_allclose(rb["omega"], W0, rtol=0, atol=0)
