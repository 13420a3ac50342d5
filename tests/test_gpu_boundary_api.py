"""The reference's array and attribute surface at the boundary, on the GPU
against the oracle:

* Solver.particle_neighbors in the reference layout [max_n][K] float4
  (x_i - x_j, bits of j) (make_neighbor_list, runner.hpp:1896-1899): ids and
  displacements bit for bit;
* particle_boundary / particle_boundary_kernel ([B][max_n] float4,
  runner.hpp:1982-2002) for every (boundary, particle) pair, and
  particle_force / particle_torque ([B][max_n] float3) of a step, for the
  moving bodies and, once recorded, the static container;
* the Pile's write-through x / v / q / omega views (pyalluvion.cu:467-470),
  its mass-property lists, get_matrix, calculate_cfl_v2, compute_mask and
  replace (pile.hpp:423-492, 816-825, 1010-1065);
* SolverIi's per-particle arrays (pyalluvion.cu:970-995).

Displacements and ids bit for bit; records, kernels and forces within 1e-5
of their scale (positions of boundary points within 1e-6 m); IISPH arrays
within 1e-5.
"""
import numpy as np
import pytest

import alluvion_amd._alluvion as al
from alluvion_amd import scenes
from gpu_sync import REL, body_state, oracle_for, rel_err, sync

pytestmark = pytest.mark.gpu


def scaled_rigid():
    return scenes.dam_break_rigid(nx=128, ny=24, nz=32, body_resolution=32)


SETTLE = 300  # steps for the bodies to settle into the fluid (gap of h at the start)


def test_particle_neighbors_reference_layout():
    sc = scaled_rigid()
    dp, pile, solver, _, _ = scenes.build_solver(sc)
    for _ in range(3):
        solver.step()
    orc = oracle_for(sc, dp, pile.volume_grids)
    sync(dp, solver, pile, orc)
    solver.update_particle_neighbors()
    orc.compute_all_boundaries()
    orc.update_particle_neighbors()
    n, K = sc.n, sc.max_num_neighbors
    rows = dp.coat(solver.particle_neighbors).get()
    assert rows.shape == (solver.max_num_particles, K, 4)
    ids, xij, cnt = orc.lists()
    gcnt = dp.coat(solver.particle_num_neighbors).get()[:n]
    np.testing.assert_array_equal(gcnt, cnt[:n])
    mask = np.arange(K)[None, :] < cnt[:n, None]
    gid = rows[:n, :, 3].copy().view(np.uint32)
    np.testing.assert_array_equal(np.where(mask, gid, 0), np.where(mask, ids[:n], 0))
    np.testing.assert_array_equal(np.where(mask[..., None], rows[:n, :, :3], 0),
                                  np.where(mask[..., None], xij[:n], 0))
    assert not rows[:n][~mask].any()  # slots past the count are zero


def test_boundary_records_and_forces_match_oracle():
    sc = scaled_rigid()
    dp, pile, solver, _, _ = scenes.build_solver(sc)
    for _ in range(SETTLE):
        solver.step()
    solver.set_static_boundary_forces(True)
    orc = oracle_for(sc, dp, pile.volume_grids)
    sync(dp, solver, pile, orc)
    solver.step()
    orc.step()
    n, nb = sc.n, len(sc.boundaries)
    rec = dp.coat(solver.particle_boundary).get()
    ker = dp.coat(solver.particle_boundary_kernel).get()
    assert rec.shape == (nb, solver.max_num_particles, 4)
    orec, oker = orc.get("boundary"), orc.get("boundary_kernel")
    for b in range(nb):
        err_x = np.abs(rec[b, :n, :3] - orec[b, :n, :3]).max()
        err_v = rel_err(rec[b, :n, 3], orec[b, :n, 3])
        err_k = rel_err(ker[b, :n], oker[b, :n])
        print("boundary %d: x_b %.3g, volume %.3g, kernel %.3g; pairs %d" % (
            b, err_x, err_v, err_k, np.count_nonzero(orec[b, :n, 3])), flush=True)
        assert err_x < 1e-6 and err_v < REL and err_k < REL, b
        assert np.count_nonzero(orec[b, :n, 3]) > 100  # every body touches the fluid
    force = dp.coat(solver.particle_force).get()
    torque = dp.coat(solver.particle_torque).get()
    assert force.shape == (nb, solver.max_num_particles, 3)
    of_, ot = orc.get("force"), orc.get("torque")
    for b in range(nb):  # the container (b = 0) too: static forces recorded
        ef, et = rel_err(force[b, :n], of_[b, :n]), rel_err(torque[b, :n], ot[b, :n])
        print("boundary %d: force %.3g torque %.3g" % (b, ef, et), flush=True)
        assert ef < REL and et < REL, b
        assert np.abs(of_[b, :n]).max() > 0


def test_static_boundary_forces_are_zero_unless_recorded():
    sc = scaled_rigid()
    dp, pile, solver, _, _ = scenes.build_solver(sc)
    for _ in range(SETTLE):
        solver.step()
    force = dp.coat(solver.particle_force).get()
    assert not force[0].any()  # the static container
    assert np.abs(force[1:, :sc.n]).max() > 0


def test_pile_views_and_helpers():
    from scipy.spatial.transform import Rotation

    sc = scaled_rigid()
    dp, pile, solver, _, _ = scenes.build_solver(sc)
    nb = len(sc.boundaries)
    assert len(pile.x) == nb and len(list(pile.omega)) == nb
    pile.x[1] = dp.f3(0.1, -0.3, -0.4)
    assert tuple(pile.get_x(1)) == tuple(np.float32([0.1, -0.3, -0.4]))
    assert tuple(pile.x[-4]) == tuple(pile.get_x(1))
    q = np.array([0.2, -0.4, 0.1, 0.9])
    q /= np.linalg.norm(q)
    pile.q[2] = dp.f4(*q)  # body 2: a cube
    pile.v[2] = dp.f3(0.0, 0.5, -0.25)
    pile.omega[4] = dp.f3(3.0, -1.0, 2.0)
    pile.v = [dp.f3(0, 0, 0)] * 2 + [pile.v[2]] + [dp.f3(0, 0, 0)] * 2
    assert tuple(pile.get_v(2)) == (0.0, 0.5, -0.25)
    with pytest.raises(IndexError):
        pile.x[nb]
    masses = pile.mass
    assert masses[1] == pytest.approx(sc.boundaries[1].mass, rel=1e-6)
    pile.friction = [0.1 * k for k in range(nb)]
    assert pile.friction[3] == pytest.approx(0.3)
    pile.mass = masses
    # get_matrix: rotation of q and the translation, glm's layout as rows
    M = np.asarray(pile.get_matrix(2))
    qs = np.array(list(pile.get_q(2)), np.float64)
    np.testing.assert_allclose(M[:3, :3], Rotation.from_quat(qs).as_matrix(), atol=2e-6)
    np.testing.assert_allclose(M[:3, 3], list(pile.get_x(2)), rtol=1e-7)
    np.testing.assert_array_equal(M[3], [0, 0, 0, 1])
    # calculate_cfl_v2: max over bodies of |omega x (max_distance, 0, 0) + v|^2
    want = 0.0
    for b, spec in enumerate(sc.boundaries):
        if spec.kind == "sphere":
            dmax = spec.params[0]
        else:
            dmax = np.linalg.norm(0.5 * np.array(spec.params[:3]))
        w = np.array(list(pile.get_omega(b)), np.float64)
        v = np.array(list(pile.get_v(b)), np.float64)
        want = max(want, float(np.sum((np.cross(w, [dmax, 0, 0]) + v) ** 2)))
    assert pile.calculate_cfl_v2() == pytest.approx(want, rel=1e-5)
    # compute_mask for the rotated cube (body 2): sign * d >= threshold
    rng = np.random.default_rng(11)
    c = np.array(list(pile.get_x(2)), np.float64)
    pts = (c + rng.uniform(-0.08, 0.08, (4096, 3))).astype(np.float32)
    sx = dp.create_coated((4096,), 3)
    sx.set(pts)
    mask = dp.create_coated((4096,), 1)
    mask.set(np.ones(4096, np.float32))
    thr = 0.01
    pile.compute_mask(2, thr, sx, mask, 4096)
    local = Rotation.from_quat(qs).inv().apply(pts.astype(np.float64) - c)
    assert sc.boundaries[2].kind == "box"
    box = scenes.BoundarySpec("box", sc.boundaries[2].params, 1.0, (1, 1, 1))
    d = scenes.body_distance(local, box)
    got = mask.get()
    clear = np.abs(d - thr) > 1e-5
    np.testing.assert_array_equal(got[clear], (d >= thr)[clear].astype(np.float32))
    assert 0 < got.sum() < 4096


def test_replace_equals_a_pile_built_with_the_new_body():
    sc = scaled_rigid()
    dp, pile, solver, _, _ = scenes.build_solver(sc)
    new_x = (-0.42, -0.29, -0.40)  # above the fluid
    q = np.array([0.0, 0.38268343, 0.0, 0.92387953])
    pile.set_v(1, dp.f3(1, 2, 3))
    pile.replace(1, dp.BoxDistance.create(dp.f3(0.06, 0.06, 0.06), 0.0),
                 al.uint3(32, 32, 32), sign=1.0, mass=0.2, inertia_tensor=dp.f3(1e-4, 1e-4, 1e-4),
                 x=dp.f3(*new_x), q=dp.f4(*q))
    assert tuple(pile.get_v(1)) == (0.0, 0.0, 0.0)
    sc2 = scaled_rigid()
    sc2.boundaries[1] = scenes.BoundarySpec("box", (0.06, 0.06, 0.06, 0.0), 1.0, (32, 32, 32),
                                            mass=0.2, inertia=(1e-4, 1e-4, 1e-4), x=new_x,
                                            q=tuple(q))
    dp2, pile2, solver2, _, _ = scenes.build_solver(sc2)
    np.testing.assert_array_equal(dp.coat(pile.volume_grids[1]).get(),
                                  dp2.coat(pile2.volume_grids[1]).get())
    for _ in range(3):
        solver.step()
        solver2.step()
    np.testing.assert_array_equal(dp.coat(solver.particle_x).get(),
                                  dp2.coat(solver2.particle_x).get())
    np.testing.assert_array_equal(dp.coat(solver.particle_v).get(),
                                  dp2.coat(solver2.particle_v).get())
    for k in ("x", "v", "q", "omega"):
        np.testing.assert_array_equal(body_state(pile, 1)[k], body_state(pile2, 1)[k])


def test_iisph_arrays_match_oracle():
    sc = scenes.dam_break_small()
    dp, pile, solver, _, _ = scenes.build_solver(sc, "ii")
    for _ in range(20):
        solver.step()
    orc = oracle_for(sc, dp, pile.volume_grids, kind="ii")
    sync(dp, solver, pile, orc, kind="ii")
    solver.step()
    orc.step()
    n = sc.n
    for name in ("last_pressure", "aii", "dii", "dij_pj", "adv_density", "pressure_accel",
                 "density_err"):
        g = dp.coat(getattr(solver, "particle_" + name)).get()[:n]
        o = orc.get(name)[:n]
        e = rel_err(g, o)
        print("IISPH %s: %.3g" % (name, e), flush=True)
        assert e < REL, (name, e)
