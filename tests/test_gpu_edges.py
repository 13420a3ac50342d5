"""Edge cases on the GPU against the oracle: a particle outside the grid,
over-full cells and lists past max_num_neighbors (anomaly flags and the
reference's rule that the last slot keeps the latest neighbour) bit for bit;
scenes that use fewer slots than max_num_particles (a count that is not a
multiple of the 256-particle block) stepped against the oracle; an empty
solver; a free rotated box driven by the fluid under each solver."""
import numpy as np
import pytest

import oracle_ffi as of
from alluvion_amd import scenes
from gpu_sync import sync
from test_gpu_parity import REL, neighbor_parity, oracle_like, rel_err

pytestmark = pytest.mark.gpu


def test_anomalies_and_capped_lists_match_oracle():
    sc = scenes.dam_break_small()
    x = sc.x.copy()
    x[0] = [5.0, 5.0, 5.0]  # outside the grid
    sc.x = x
    sc.max_num_particles_per_cell = 4  # lattice cells hold 8
    sc.max_num_neighbors = 16  # interior particles have about 33
    dp, solver, orc = neighbor_parity(sc)
    _, _, anomaly = orc.grid()
    assert anomaly.tolist() == [1, 1, 1]
    _, cnt = solver.get_neighbors()
    assert cnt.max() == 16


@pytest.mark.parametrize("kind", ["df", "i"])
def test_fewer_particles_than_slots(kind):
    sc = scenes.dam_break_small()
    keep = sc.n - 1037  # not a multiple of 256
    sc.x, sc.v = sc.x[:keep], sc.v[:keep]
    cap = 16384 + 300
    dp, pile, solver, _, _ = scenes.build_solver(sc, kind, max_particles=cap)
    orc = of.Oracle(sc, precision=4, kind=kind, max_particles=cap)
    for b, var in enumerate(pile.volume_grids):
        orc.set_volume_map(b, dp.coat(var).get())
    assert solver.max_num_particles == cap and solver.num_particles == keep
    for step in range(3):
        if step:  # each step from the GPU's state
            sync(dp, solver, pile, orc, kind=kind)
        solver.step()
        orc.step()
        assert solver.num_density_solve == orc.report()["num_density_solve"], step
        x = dp.coat(solver.particle_x).get()[:keep]
        v = dp.coat(solver.particle_v).get()[:keep]
        assert np.abs(x - orc.get("x")[:keep]).max() < 1e-6, step
        assert rel_err(v, orc.get("v")[:keep]) < REL, step


def test_empty_solver_steps():
    """No particles: the step launches no particle kernel, the grid stays
    empty and time advances by dt as in the oracle.  The oracle, like the
    reference, runs the minimum solve counts on a NaN mean (0/0); the HIP
    path runs no iteration and reports zero counts."""
    sc = scenes.dam_break_small()
    dp, pile, solver, _, _ = scenes.build_solver(sc, "df")
    orc = oracle_like(sc, dp, pile)
    solver.num_particles = 0
    orc.L.orc_set_num_particles(orc.h, 0)
    for _ in range(2):
        solver.step()
        orc.step()
    assert solver.t == pytest.approx(orc.report()["t"], rel=1e-6)
    assert solver.num_divergence_solve == 0 and solver.num_density_solve == 0
    solver.update_particle_neighbors()
    count, _, anomaly = solver.get_grid()
    assert not count.any() and not anomaly.any()


@pytest.mark.parametrize("kind", ["df", "ii", "i"])
def test_free_body_driven_by_fluid_matches_oracle(kind):
    """A rotated free box hit by a fluid block (no gravity, no viscosity):
    fluid velocities and the body's kinematics over three steps."""
    x = scenes.lattice_block(12, 12, 12, 0.0025, -0.03, 7, 0.00025)
    x[:, 1] += np.float32(0.044)
    v = np.zeros_like(x)
    v[:, 0], v[:, 1] = 0.2, -1.0
    q = np.array([0.1, 0.3, -0.2, 0.9])
    q /= np.linalg.norm(q)
    body = scenes.BoundarySpec("box", (0.03, 0.02, 0.025, 0.0), 1.0, (16, 16, 16), mass=0.05,
                               inertia=(2e-5, 3e-5, 4e-5), x=(0.004, 0.0, 0.003), q=tuple(q))
    sc = scenes.Scene(name="fluid_on_free_box", x=x.astype(np.float32), v=v.astype(np.float32),
                      gravity=(0, 0, 0), pile_gravity=(0, 0, 0), viscosity=0.0,
                      boundary_viscosity=0.0, boundaries=[body], grid_offset=(-8, -8, -8),
                      grid_res=(16, 16, 16))
    dp, pile, solver, _, _ = scenes.build_solver(sc, kind)
    orc = oracle_like(sc, dp, pile, kind=kind)
    n = sc.n
    for step in range(3):
        if step:  # each step from the GPU's state, the body's included
            sync(dp, solver, pile, orc, kind=kind)
        solver.step()
        orc.step()
        assert solver.num_density_solve == orc.report()["num_density_solve"], step
        assert rel_err(dp.coat(solver.particle_v).get()[:n], orc.get("v")[:n]) < REL, step
        rb = orc.rigid(0)
        qg = pile.get_q(0)
        np.testing.assert_allclose(list(pile.get_x(0)), rb["x"], rtol=1e-6, atol=1e-8, err_msg=str(step))
        np.testing.assert_allclose(list(pile.get_v(0)), rb["v"], rtol=1e-4, atol=1e-6,
                                   err_msg=str(step))
        np.testing.assert_allclose(list(pile.get_omega(0)), rb["omega"], rtol=1e-4,
                                   atol=1e-5, err_msg=str(step))
        np.testing.assert_allclose([qg.x, qg.y, qg.z, qg.w], rb["q"], rtol=0, atol=1e-6,
                                   err_msg=str(step))
    assert np.abs(orc.rigid(0)["omega"]).max() > 1e-3  # the fluid turns the box
