"""Pins the oracle's viscosity (compute_viscosity, runner.hpp:2150-2249)
with facts that do not go through its formula:

* on a perfect periodic lattice, for interior particles, a uniform and a
  linear velocity field give no viscous acceleration (the pair terms are
  radial weights times v_i - v_j, whose lattice sums vanish by symmetry);
* a quadratic field v = (x^2, 0, 0) gives nu * (1/2) Laplacian(v) = nu in x
  (the continuum limit of sum_j V_j (x_ij . grad W) / |x_ij|^2 (v_i - v_j)
  is (1/2) Laplacian(v) for a normalised kernel) to 2 %, the gap being the
  0.01 h^2 regulariser and the lattice; nothing in y and z;
* Galilean invariance: adding one velocity to every particle and to the
  boundaries (a moving container) changes no acceleration, fluid or
  boundary term.
"""
import numpy as np

import oracle_ffi as of
from alluvion_amd import scenes
from test_oracle_solver import lattice_sum

NU = 0.01


# binary-exact lattice (r = 2^-9, h = 2^-7): FP32 positions keep the
# lattice's symmetry exactly
RB, HB = 2.0 ** -9, 2.0 ** -7


def lattice_accel(vfun, m=24):
    d = 2 * RB
    L = m * d
    g = np.arange(m) * d - 0.5 * L + 0.5 * d
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    S = lattice_sum(d, HB)
    c = int(round(L / HB))
    sc = scenes.Scene(name="viscosity_lattice", x=x.astype(np.float32),
                      v=np.zeros_like(x, np.float32), particle_radius=RB, kernel_radius=HB,
                      particle_mass=1000.0 / S, wrap_length=(L, L, L), gravity=(0, 0, 0),
                      viscosity=NU, boundary_viscosity=NU, grid_offset=(-c // 2,) * 3,
                      grid_res=(c,) * 3, boundaries=[], pile_gravity=(0, 0, 0))
    orc = of.Oracle(sc, precision=8, kind="df")
    orc.update_settings(enable_divergence_solve=0, enable_density_solve=0, dt=1e-9,
                        max_dt=1e-9, min_dt=1e-9)
    orc.set("v", vfun(x))
    orc.step()
    inner = np.all(np.abs(x) < 0.5 * L - 2 * HB, axis=1)
    assert inner.sum() >= 512
    return orc.get("a")[:sc.n][inner]


def test_uniform_and_linear_fields_have_no_viscous_acceleration():
    a = lattice_accel(lambda x: np.tile([0.3, -0.7, 0.2], (x.shape[0], 1)))
    assert np.abs(a).max() < 1e-15
    G = np.array([[0.5, -1.0, 2.0], [1.5, 0.2, -0.3], [-0.8, 0.6, 1.1]])
    a = lattice_accel(lambda x: x @ G.T)
    assert np.abs(a).max() < 1e-12


def test_quadratic_field_gives_half_the_laplacian():
    def quad(x):
        v = np.zeros_like(x)
        v[:, 0] = x[:, 0] ** 2
        return v
    a = lattice_accel(quad)
    np.testing.assert_allclose(a[:, 0], NU, rtol=0.02)
    assert np.abs(a[:, 1:]).max() < 1e-15
    # Galilean invariance on the lattice
    a2 = lattice_accel(lambda x: quad(x) + np.array([0.4, -0.1, 0.25]))
    np.testing.assert_allclose(a2, a, rtol=0, atol=1e-12 * NU)


def test_moving_container_and_fluid_together_feel_no_viscosity():
    # dam break in a container: every fluid particle and the container move
    # with one velocity U; both the fluid and the boundary terms vanish
    sc = scenes.dam_break_small()
    sc.boundaries[0].resolution = (8, 8, 8)
    sc.gravity = (0.0, 0.0, 0.0)
    U = np.array([0.3, -0.2, 0.5])
    orc = of.Oracle(sc, precision=8, kind="df")
    orc.build_volume_map(0)
    orc.update_settings(enable_divergence_solve=0, enable_density_solve=0, dt=1e-9,
                        max_dt=1e-9, min_dt=1e-9)
    n = sc.n
    v = np.zeros((orc.nt, 3))
    v[:n] = U
    orc.set("v", v)
    orc.set_rigid(0, v=list(U))
    orc.step()
    a_moving = orc.get("a")[:n]
    assert np.abs(a_moving).max() < 1e-12
    # the container at rest under the same fluid: the boundary term acts
    orc2 = of.Oracle(sc, precision=8, kind="df")
    orc2.build_volume_map(0)
    orc2.update_settings(enable_divergence_solve=0, enable_density_solve=0, dt=1e-9,
                         max_dt=1e-9, min_dt=1e-9)
    orc2.set("v", v)
    orc2.step()
    a_rest = orc2.get("a")[:n]
    wall = np.abs(a_rest).max(1) > 0
    assert wall.sum() > 100  # the particles next to the walls are braked
    assert np.all((a_rest[wall] @ U) < 0)
