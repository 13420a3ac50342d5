"""Particle-sampled boundaries (pellets) on the GPU against the oracle:
the fluid and pellet lists bit for bit, the density with pellets, ISPH steps
with pellets (FP32, 1e-5 on the first step), and the pellet contacts."""
import numpy as np
import pytest

import oracle_ffi as of
from test_gpu_contacts import compare_contacts
from gpu_sync import sync
from test_oracle_pellets import small_scene

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


def pellet_lists_orc(orc, pellet_list, rows):
    ids, _, cnt = orc.lists(pellet_list)
    return ids[rows], cnt[rows]


def test_pellet_lists_and_density_match_oracle():
    from alluvion_amd import scenes
    sc = small_scene()
    dp, pile, solver, _, _ = scenes.build_solver(sc, kind="i")
    assert pile.num_pellets == sum(b.collision_vertices.shape[0] for b in sc.boundaries)
    solver.transform_all_pellets()
    solver.update_particle_neighbors()
    orc = of.Oracle(sc, precision=4, kind="i")
    orc.transform_pellets()
    orc.update_particle_neighbors()
    n, mx, npel = sc.n, orc.max_n, pile.num_pellets
    K = sc.max_num_neighbors
    x_g = dp.coat(solver.particle_x).get()
    np.testing.assert_array_equal(x_g[mx:mx + npel], orc.get("x")[mx:mx + npel])
    ids, cnt = solver.get_neighbors()
    oids, ocnt = pellet_lists_orc(orc, False, slice(0, n))
    np.testing.assert_array_equal(cnt[:n], ocnt)
    for i in range(n):
        np.testing.assert_array_equal(ids[i, :cnt[i]], oids[i, :cnt[i]], err_msg=str(i))
    pids, pcnt = solver.get_pellet_neighbors(0, mx)
    oids, ocnt = pellet_lists_orc(orc, True, slice(0, n))
    np.testing.assert_array_equal(pcnt[:n], ocnt)
    assert ocnt.max() > 0
    for i in range(n):
        np.testing.assert_array_equal(pids[i, :pcnt[i]], oids[i, :pcnt[i]], err_msg=str(i))
    bids, bcnt = solver.get_pellet_neighbors(1, sc.max_num_pellets)
    oids, ocnt = pellet_lists_orc(orc, False, slice(mx, mx + npel))
    np.testing.assert_array_equal(bcnt[:npel], ocnt)
    assert ocnt.max() > 0
    for k in range(npel):
        np.testing.assert_array_equal(bids[k, :bcnt[k]], oids[k, :bcnt[k]], err_msg=str(k))
    assert K == ids.shape[1]
    dens = dp.coat(solver.particle_density).get()[:n]
    assert rel(dens, orc.get("density")[:n]) < 1e-6
    assert np.all(dp.coat(solver.grid_anomaly).get() == 0)


def body_state(pile, b):
    q = pile.get_q(b)
    return (np.array(list(pile.get_x(b))), np.array(list(pile.get_v(b))),
            np.array([q.x, q.y, q.z, q.w]), np.array(list(pile.get_omega(b))))


def test_isph_steps_with_pellets_match_oracle():
    from alluvion_amd import scenes
    sc = small_scene()
    dp, pile, solver, _, _ = scenes.build_solver(sc, kind="i")
    orc = of.Oracle(sc, precision=4, kind="i")
    n, mx, npel = sc.n, orc.max_n, pile.num_pellets
    for step in range(5):
        if step:  # each step from the GPU's state, bodies and pellets included
            sync(dp, solver, pile, orc, kind="i")
        solver.step()
        orc.step()
        x_g = dp.coat(solver.particle_x).get()
        v_g = dp.coat(solver.particle_v).get()
        p_g = dp.coat(solver.particle_pressure).get()
        assert rel(x_g[:n], orc.get("x")[:n]) < 1e-6, step
        assert rel(v_g[:n], orc.get("v")[:n]) < 1e-5, step
        p_o = orc.get("pressure")
        # the pressure is a Jacobi solution: one FP32 rounding per sweep
        assert rel(p_g[:n], p_o[:n]) < 1e-4, step
        assert rel(p_g[mx:mx + npel], p_o[mx:mx + npel]) < 1e-4, step
        assert solver.num_density_solve == orc.report()["num_density_solve"], step
        x, v, q, w = body_state(pile, 1)
        rb = orc.rigid(1)
        np.testing.assert_allclose(x, rb["x"], rtol=1e-6, atol=1e-8, err_msg=str(step))
        np.testing.assert_allclose(v, rb["v"], rtol=1e-4, atol=1e-6, err_msg=str(step))
        np.testing.assert_allclose(w, rb["omega"], rtol=1e-4, atol=1e-5, err_msg=str(step))
    assert np.abs(orc.rigid(1)["v"]).max() > 0  # the fluid pushes the sphere
    assert p_o[mx:mx + npel].max() > 0  # pellets carry pressure


def test_pellet_contacts_match_oracle():
    from alluvion_amd import scenes
    sc = small_scene()
    sc.boundaries[1].x = (0.01, -0.07 + 0.016, 0.0)  # into the floor
    sc.contact_tolerance = 1e-4
    dp, pile, solver, _, _ = scenes.build_solver(sc, kind="i")
    pile.set_v(1, dp.f3(0.3, -0.8, 0.1))
    pile.set_omega(1, dp.f3(1.0, 0.5, -2.0))
    orc = of.Oracle(sc, precision=4, kind="i")
    orc.set_rigid(1, v=[0.3, -0.8, 0.1], omega=[1.0, 0.5, -2.0])
    solver.transform_all_pellets()
    pile.copy_kinematics_to_device()
    n = pile.find_contacts(solver.particle_x, solver.max_num_particles)
    n_orc = orc.find_contacts_pellets()
    assert n == n_orc > 5, (n, n_orc)
    compare_contacts(pile.get_contacts(), of.contact_dicts(orc.contacts()))
    pile.solve_contacts()
    orc.solve_contacts()
    x, v, q, w = body_state(pile, 1)
    rb = orc.rigid(1)
    np.testing.assert_allclose(v, rb["v"], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(w, rb["omega"], rtol=1e-5, atol=1e-5)


def test_runner_pellet_lists_and_sums_match_oracle():
    from alluvion_amd import scenes
    sc = small_scene()
    dp, pile, solver, _, _ = scenes.build_solver(sc, kind="i")
    solver.transform_all_pellets()
    solver.update_particle_neighbors()
    orc = of.Oracle(sc, precision=4, kind="i")
    orc.transform_pellets()
    orc.update_particle_neighbors()
    runner = dp.Runner()
    mx, K = solver.max_num_particles, sc.max_num_neighbors
    off, ns = 100, 60
    rows = off + ns
    bead, pel = dp.create((rows, K), 4), dp.create((rows, K), 4)
    nb, npl = dp.create((rows,), 1, np.uint32), dp.create((rows,), 1, np.uint32)
    for c in (nb, npl):
        dp.coat(c).set(np.zeros(rows, np.uint32))
    runner.launch_make_bead_pellet_neighbor_list(solver.particle_x, solver.pid, solver.pid_length,
                                                 bead, nb, pel, npl, solver.grid_anomaly, mx,
                                                 ns, off)
    bid, bxx, bc, pid, pxx, pc = orc.sample_bead_pellet(orc.get("x")[:rows], offset=off)
    g_b = dp.coat(bead).get().reshape(rows, K, 4)
    g_p = dp.coat(pel).get().reshape(rows, K, 4)
    np.testing.assert_array_equal(dp.coat(nb).get(), bc)
    np.testing.assert_array_equal(dp.coat(npl).get(), pc)
    assert pc[off:].max() > 0
    for i in range(off, rows):
        np.testing.assert_array_equal(g_b[i, :bc[i], 3].view(np.uint32), bid[i, :bc[i]])
        np.testing.assert_array_equal(g_p[i, :pc[i], 3].view(np.uint32), pid[i, :pc[i]])
        np.testing.assert_allclose(g_b[i, :bc[i], :3], bxx[i, :bc[i]], atol=1e-9)
        np.testing.assert_allclose(g_p[i, :pc[i], :3], pxx[i, :pc[i]], atol=1e-9)
    # sums over the GPU lists against the oracle's sums over its lists
    dens, fac = dp.create((rows,), 1), dp.create((rows,), 1)
    bk = dp.create((rows,), 4)
    runner.launch_compute_density_with_pellets(bead, nb, pel, npl, dens, rows)
    runner.launch_compute_dfsph_factor_with_pellets(bead, nb, pel, npl, fac, rows)
    runner.launch_compute_particle_boundary_with_pellets(bk, pel, npl, rows)
    s = slice(off, rows)
    assert rel(dp.coat(dens).get()[s], orc.pellet_sums(1, pxx[s], pc[s], bxx[s], bc[s])) < 1e-6
    assert rel(dp.coat(fac).get()[s], orc.pellet_sums(2, pxx[s], pc[s], bxx[s], bc[s])) < 1e-5
    assert rel(dp.coat(bk).get().reshape(rows, 4)[s], orc.pellet_sums(0, pxx[s], pc[s])) < 1e-5
    # sample lists of a pellet solver hold the pellets too
    sx = dp.create_coated((ns,), 3)
    pts = orc.get("x")[off:rows].astype(np.float32)
    sx.set(pts)
    nbr, cnt = dp.create((ns, K), 4), dp.create((ns,), 1, np.uint32)
    runner.launch_make_neighbor_list(sx, solver.pid, solver.pid_length, nbr, cnt,
                                     solver.grid_anomaly, ns)
    ids_o, xixj_o, cnt_o = orc.sample_neighbors(pts)
    np.testing.assert_array_equal(dp.coat(cnt).get(), cnt_o)
    lists = dp.coat(nbr).get().reshape(ns, K, 4)
    assert (ids_o >= mx).any()
    for k in range(ns):
        np.testing.assert_array_equal(lists[k, :cnt_o[k], 3].view(np.uint32), ids_o[k, :cnt_o[k]])


def test_sampling_with_pellets_matches_oracle():
    from alluvion_amd import scenes
    sc = small_scene()
    dp, pile, solver, _, _ = scenes.build_solver(sc, kind="i")
    for _ in range(3):
        solver.step()
    solver.update_particle_neighbors()  # grid of the current positions
    orc = of.Oracle(sc, precision=4, kind="i")
    runner = dp.Runner()
    mx, K = solver.max_num_particles, sc.max_num_neighbors
    x = dp.coat(solver.particle_x).get()
    rng = np.random.default_rng(4)
    ns = 200
    pts = (x[rng.choice(sc.n, ns)] + rng.uniform(-0.004, 0.004, (ns, 3))).astype(np.float32)
    sx = dp.create_coated((ns,), 3)
    sx.set(pts)
    bead, pel = dp.create((ns, K), 4), dp.create((ns, K), 4)
    nb, npl = dp.create((ns,), 1, np.uint32), dp.create((ns,), 1, np.uint32)
    runner.launch_make_bead_pellet_neighbor_list(sx, solver.pid, solver.pid_length, bead, nb, pel,
                                                 npl, solver.grid_anomaly, mx, ns, 0)
    g_b = dp.coat(bead).get().reshape(ns, K, 4)
    g_p = dp.coat(pel).get().reshape(ns, K, 4)
    lists = (np.ascontiguousarray(g_b[..., 3]).view(np.uint32), g_b[..., :3].astype(np.float64),
             dp.coat(nb).get())
    pl = (np.ascontiguousarray(g_p[..., 3]).view(np.uint32), g_p[..., :3].astype(np.float64),
          dp.coat(npl).get())
    assert pl[2].max() > 0
    dens = dp.coat(solver.particle_density).get().astype(np.float64)
    v = dp.coat(solver.particle_v).get().astype(np.float64)
    assert np.abs(v[mx:mx + pile.num_pellets]).max() > 0  # the sphere's pellets move
    out1 = dp.create((ns,), 1)
    runner.launch_sample_density_with_pellets(sx, bead, nb, pel, npl, out1, ns)
    assert rel(dp.coat(out1).get(), orc.sample_with_pellets(0, lists, pl)) < 1e-6
    sv = dp.create((ns,), 3)
    runner.launch_sample_velocity_with_pellets(sx, solver.particle_x, solver.particle_density,
                                               solver.particle_v, bead, nb, sv, pel, npl, ns)
    sv_o = orc.sample_with_pellets(1, lists, pl, dens, v)
    assert rel(dp.coat(sv).get(), sv_o) < 1e-5
    vort = dp.create((ns,), 3)
    runner.launch_sample_vorticity_with_pellets(sx, solver.particle_x, solver.particle_density,
                                                solver.particle_v, bead, nb, sv, vort, pel, npl,
                                                ns)
    svg = dp.coat(sv).get().astype(np.float64)
    assert rel(dp.coat(vort).get(), orc.sample_with_pellets(2, lists, pl, dens, v, sample_v=svg)) < 1e-4
