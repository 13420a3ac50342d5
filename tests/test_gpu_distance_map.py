"""Boundaries whose signed distance is sampled on their map nodes (the
reference's grid path, which it takes for MeshDistance: pile.hpp:79-95,
508-535; compute_volume_and_boundary_x and compute_particle_boundary,
runner.hpp:1696-1736, 2005-2025; update_volume_field from a distance map,
runner.hpp:1616-1658; collision_find_dist_normal, runner.hpp:1794-1825), on
the GPU against the oracle's restatement of the same path, fed the GPU's
distance maps:

* the distance grid holds the sampled distance at every node, and analytic
  boundaries have an empty one (Pile.distance_grids, pile.hpp:981-988);
* the volume map built from the distance map, within 1e-5 of the ball
  volume;
* boundary records and kernels of a floating-body dam break whose container
  and four bodies are all sampled, after 300 settling steps (positions of
  boundary points within 1e-6 m, volumes and kernels within 1e-5), and the
  following steps (fields within 1e-5, equal iteration counts, body forces
  and kinematics as in the analytic-boundary tests);
* contacts of a rotated sphere's vertices against a sampled ground.
"""
import numpy as np
import pytest

import oracle_ffi as of
from alluvion_amd import scenes
from gpu_sync import REL, body_state, compare_fields, oracle_for, rel_err, sync
from test_gpu_configs import check_bodies, step_pair
from test_gpu_contacts import compare_contacts, rotated_sphere_scene

pytestmark = pytest.mark.gpu

SETTLE = 300


def sampled(sc):
    for b in sc.boundaries:
        b.sampled = True
    return sc


def test_distance_grids_hold_the_sampled_distance():
    sc = scenes.dam_break_small()
    sc.boundaries[0].resolution = (24, 24, 24)
    sc.boundaries.append(scenes.BoundarySpec("sphere", (0.05,), 1.0, (12, 12, 12),
                                             x=(0.1, 0.1, 0.1), sampled=True))
    dp, pile, solver, _, _ = scenes.build_solver(sc)
    grids = pile.distance_grids
    assert dp.coat(grids[0]).get().size == 0  # analytic: empty, as the reference's
    got = dp.coat(grids[1]).get()
    orc = of.Oracle(sc, precision=8)
    dmin, dmax, res, cs = orc.grid_info(1)
    assert got.size == orc.num_nodes(1)
    want = scenes.shape_distance("sphere", (0.05,), of.node_positions(dmin, res, cs))
    # node positions are computed in FP32 by the library, in FP64 here
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-7)
    # compute_mask against the sampled sphere: sign * (interpolated map) >=
    # threshold, the map interpolated by the oracle from the GPU's nodes;
    # points off the map have no distance (FLT_MAX) and keep their mask
    rng = np.random.default_rng(9)
    sx = (np.array([0.1, 0.1, 0.1]) + rng.uniform(-0.09, 0.09, (4096, 3))).astype(np.float32)
    thr = 0.004
    mask_var = dp.create_coated((4096,), 1, np.float32)
    mask_var.set(np.ones(4096, np.float32))
    sx_var = dp.create_coated((4096,), 3, np.float32)
    sx_var.set(sx)
    pile.compute_mask(1, thr, sx_var, mask_var, 4096)
    mask = mask_var.get()
    val, _, inside = of.map_value_gradient(dmin, dmax, res, cs, got,
                                           sx.astype(np.float64) - np.array([0.1, 0.1, 0.1]))
    d = np.where(inside, val, np.finfo(np.float32).max)
    clear = np.abs(d - thr) > 1e-5
    np.testing.assert_array_equal(mask[clear], (d[clear] >= thr).astype(np.float32))
    assert (mask == 0).sum() > 100 and (mask == 1).sum() > 100


def test_volume_map_from_the_distance_map_matches_oracle():
    sc = sampled(scenes.dam_break_small())
    sc.boundaries[0].resolution = (24, 24, 24)
    dp, pile, solver, _, _ = scenes.build_solver(sc)
    orc = oracle_for(sc, dp, [], distance_grids=pile.distance_grids)
    orc.build_volume_map(0)
    got = dp.coat(pile.volume_grids[0]).get()
    want = orc.volume_map(0)
    ball = 4.0 / 3.0 * np.pi * sc.kernel_radius ** 3
    err = np.abs(got - want).max() / ball
    print("volume map from distance: %.3g of the ball volume, %d nodes" % (err, got.size),
          flush=True)
    assert err < 1e-5
    assert (want > 0).sum() > 1000 and (want == 0).sum() > 1000


def test_records_and_steps_with_sampled_boundaries_match_oracle():
    sc = sampled(scenes.dam_break_rigid(nx=128, ny=24, nz=32, body_resolution=32))
    dp, pile, solver, _, _ = scenes.build_solver(sc)
    for _ in range(SETTLE):
        solver.step()
    orc = oracle_for(sc, dp, pile.volume_grids, distance_grids=pile.distance_grids)
    forces = []
    for step in range(3):
        sync(dp, solver, pile, orc)
        before, rep, _ = step_pair(dp, solver, pile, orc)
        if step == 0:
            n, nb = sc.n, len(sc.boundaries)
            rec = dp.coat(solver.particle_boundary).get()
            ker = dp.coat(solver.particle_boundary_kernel).get()
            orec, oker = orc.get("boundary"), orc.get("boundary_kernel")
            for b in range(nb):
                err_x = np.abs(rec[b, :n, :3] - orec[b, :n, :3]).max()
                err_v = rel_err(rec[b, :n, 3], orec[b, :n, 3])
                err_k = rel_err(ker[b, :n], oker[b, :n])
                pairs = np.count_nonzero(orec[b, :n, 3])
                print("sampled boundary %d: x_b %.3g, volume %.3g, kernel %.3g; pairs %d" % (
                    b, err_x, err_v, err_k, pairs), flush=True)
                assert err_x < 1e-6 and err_v < REL and err_k < REL, b
                assert pairs > 100, b
        compare_fields(dp, solver, orc, tag="sampled boundaries step %d" % step)
        forces.append(check_bodies(pile, orc, before, float(solver.dt), "step %d" % step))
    assert min(forces) > 0.1


def test_contacts_against_a_sampled_ground_match_oracle():
    sc = rotated_sphere_scene()
    sc.boundaries[0].sampled = True
    sc.boundaries[0].resolution = (16, 8, 16)
    dp, pile, solver, _, _ = scenes.build_solver(sc)
    pile.set_v(1, dp.f3(0.4, -1.1, 0.25))
    pile.set_omega(1, dp.f3(1.5, -0.5, 2.0))
    orc = oracle_for(sc, dp, pile.volume_grids, distance_grids=pile.distance_grids)
    orc.set_rigid(1, v=[0.4, -1.1, 0.25], omega=[1.5, -0.5, 2.0])
    n = pile.find_contacts()
    n_orc = orc.find_contacts()
    assert n == n_orc > 10, (n, n_orc)
    compare_contacts(pile.get_contacts(), of.contact_dicts(orc.contacts()))
    pile.solve_contacts()
    orc.solve_contacts()
    compare_contacts(pile.get_contacts(), of.contact_dicts(orc.contacts()))
