"""Edge cases on the GPU against the oracle: a particle outside the grid,
over-full cells and lists past max_num_neighbors (anomaly flags and the
reference's rule that the last slot keeps the latest neighbour) bit for bit;
scenes that use fewer slots than max_num_particles (a count that is not a
multiple of the 256-particle block) stepped against the oracle; an empty
solver."""
import numpy as np
import pytest

import oracle_ffi as of
from alluvion_amd import scenes
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
        solver.step()
        orc.step()
        assert solver.num_density_solve == orc.report()["num_density_solve"], step
        x = dp.coat(solver.particle_x).get()[:keep]
        v = dp.coat(solver.particle_v).get()[:keep]
        assert np.abs(x - orc.get("x")[:keep]).max() < 1e-6, step
        assert rel_err(v, orc.get("v")[:keep]) < REL * 10 ** step, step


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
