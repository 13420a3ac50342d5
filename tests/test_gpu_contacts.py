"""Rigid contacts on the GPU against the oracle (FP32): the contact records
in (i, j, vertex) order, the impulse solve, and whole steps with a body
resting on the ground."""
import numpy as np
import pytest

import oracle_ffi as of
from alluvion_amd import scenes
from test_oracle_contacts import TOP, contact_scene, f32, uv_sphere

pytestmark = pytest.mark.gpu

VECS = of.CONTACT_VECTORS
SCALARS = of.CONTACT_SCALARS


def rotated_sphere_scene(tol=1e-3, friction=0.3):
    q = np.array([0.2, 0.1, -0.3, 0.9])
    q /= np.linalg.norm(q)
    return contact_scene(f32(uv_sphere(0.05, 24, 12)), (0.05, TOP + 0.035, -0.02), q,
                         restitution=0.6, friction=friction, tol=tol)


def compare_contacts(got, want):
    assert len(got) == len(want) > 0, (len(got), len(want))
    for k, (g, w) in enumerate(zip(got, want)):
        assert (g["i"], g["j"]) == (w["i"], w["j"]), (k, g, w)
        for name in VECS:
            np.testing.assert_allclose(g[name], w[name], rtol=1e-5, atol=2e-6,
                                       err_msg="%s of contact %d" % (name, k))
        for name in SCALARS:
            np.testing.assert_allclose(g[name], w[name], rtol=1e-5, atol=1e-6,
                                       err_msg="%s of contact %d" % (name, k))


def body_state(pile, b):
    q = pile.get_q(b)
    return (np.array(list(pile.get_x(b))), np.array(list(pile.get_v(b))),
            np.array([q.x, q.y, q.z, q.w]), np.array(list(pile.get_omega(b))))


def test_contacts_and_solve_match_oracle():
    sc = rotated_sphere_scene()
    dp, pile, solver, _, _ = scenes.build_solver(sc)
    pile.set_v(1, dp.f3(0.4, -1.1, 0.25))
    pile.set_omega(1, dp.f3(1.5, -0.5, 2.0))
    orc = of.Oracle(sc, precision=4)
    orc.set_rigid(1, v=[0.4, -1.1, 0.25], omega=[1.5, -0.5, 2.0])
    n = pile.find_contacts()
    n_orc = orc.find_contacts()
    assert n == n_orc > 10, (n, n_orc)
    compare_contacts(pile.get_contacts(), of.contact_dicts(orc.contacts()))
    pile.solve_contacts()
    orc.solve_contacts()
    compare_contacts(pile.get_contacts(), of.contact_dicts(orc.contacts()))
    x, v, q, w = body_state(pile, 1)
    rb = orc.rigid(1)
    np.testing.assert_allclose(v, rb["v"], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(w, rb["omega"], rtol=1e-5, atol=1e-5)


def test_pair_overloads_and_capacity():
    sc = rotated_sphere_scene()
    sc.max_num_contacts = 7
    dp, pile, solver, _, _ = scenes.build_solver(sc)
    orc = of.Oracle(sc, precision=4)
    total = orc.find_contacts()
    assert total > 7, total
    n = pile.find_contacts()
    assert n == total, (n, total)
    assert len(pile.get_contacts()) == 7, len(pile.get_contacts())  # max_num_contacts kept
    compare_contacts(pile.get_contacts(), of.contact_dicts(orc.contacts()))
    assert pile.find_contacts(1, 0) == total
    assert pile.find_contacts(1) == total
    assert pile.find_contacts(0) == 0  # the ground has no collision vertices


def test_steps_with_contacts_match_oracle():
    sc = contact_scene(f32([[0.0, -0.05, 0.0], [0.03, -0.04, 0.0], [-0.02, -0.045, 0.01]]),
                       (0.0, TOP + 0.0495, 0.0), mass=0.4, inertia=(1e-3, 1e-3, 1e-3),
                       restitution=0.2, friction=0.4, tol=1e-4)
    dp, pile, solver, _, _ = scenes.build_solver(sc)
    pile.set_v(1, dp.f3(0.2, -0.5, 0.0))
    orc = of.Oracle(sc, precision=4)
    for b in range(2):
        orc.set_volume_map(b, dp.coat(pile.volume_grids[b]).get())
    orc.set_rigid(1, v=[0.2, -0.5, 0.0])
    for step in range(5):
        solver.step()
        orc.step()
        x, v, q, w = body_state(pile, 1)
        rb = orc.rigid(1)
        np.testing.assert_allclose(x, rb["x"], rtol=1e-6, atol=1e-7, err_msg=str(step))
        np.testing.assert_allclose(v, rb["v"], rtol=1e-5, atol=1e-5, err_msg=str(step))
        np.testing.assert_allclose(q, rb["q"], rtol=1e-5, atol=1e-6, err_msg=str(step))
        np.testing.assert_allclose(w, rb["omega"], rtol=1e-5, atol=1e-4, err_msg=str(step))
    assert len(pile.get_contacts()) > 0
