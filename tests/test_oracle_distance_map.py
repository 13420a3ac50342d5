"""The oracle's distance-map boundary path (the reference's grid path, which
it takes for MeshDistance): interpolate_and_derive (runner.hpp:1773-1791),
compute_volume_and_boundary_x (runner.hpp:1696-1736) and
compute_particle_boundary (runner.hpp:2005-2025).

Pinned by: exact reproduction of complete cubic polynomials, values and
world-space gradients, on an anisotropic grid (the 32-node element spans the
cubics; a transposed or mis-scaled derivative fails it); and, where a box's
signed distance is linear (near the middle of a face, away from its edges),
the map path's boundary point and volume equal the analytic path's, since
the map then interpolates the distance exactly.
"""
import itertools

import numpy as np

import oracle_ffi as of
from alluvion_amd import scenes

CUBICS = [m for m in itertools.product(range(4), repeat=3) if sum(m) <= 3]


def test_map_reproduces_cubics_with_world_gradients():
    assert len(CUBICS) == 20
    rng = np.random.default_rng(11)
    res = (3, 4, 5)
    dmin = np.array([-0.31, 0.05, -0.7])
    dmax = np.array([0.29, 0.65, 0.3])
    cs = (dmax - dmin) / np.array(res)
    pos = of.node_positions(dmin, res, cs)
    assert np.array_equal(pos[77], of.node_position(dmin, res, cs, 77))
    coef = rng.normal(size=len(CUBICS))

    def f(p):
        return sum(c * p[:, 0] ** a * p[:, 1] ** b * p[:, 2] ** e
                   for c, (a, b, e) in zip(coef, CUBICS))

    def grad(p):
        g = np.zeros_like(p)
        for c, (a, b, e) in zip(coef, CUBICS):
            x, y, z = p[:, 0], p[:, 1], p[:, 2]
            if a:
                g[:, 0] += c * a * x ** (a - 1) * y ** b * z ** e
            if b:
                g[:, 1] += c * b * x ** a * y ** (b - 1) * z ** e
            if e:
                g[:, 2] += c * e * x ** a * y ** b * z ** (e - 1)
        return g

    pts = dmin + rng.random((400, 3)) * (dmax - dmin)
    val, gr, inside = of.map_value_gradient(dmin, dmax, res, cs, f(pos), pts)
    assert inside.all()
    np.testing.assert_allclose(val, f(pts), rtol=0, atol=1e-11)
    np.testing.assert_allclose(gr, grad(pts), rtol=0, atol=1e-10)
    # outside the map: no value, no gradient
    val, gr, inside = of.map_value_gradient(dmin, dmax, res, cs, f(pos),
                                            np.array([[0.0, 0.0, 0.0], [0.3, 0.1, 0.0]]))
    assert list(inside) == [False, False] and not val.any() and not gr.any()


def test_map_record_equals_analytic_where_the_distance_is_linear():
    h, r = 0.01, 0.0025
    box = (0.2, 0.16, 0.12, 0.0)
    analytic = scenes.BoundarySpec("box", box, -1.0, (20, 16, 12))
    mapped = scenes.BoundarySpec("box", box, -1.0, (20, 16, 12), sampled=True)
    x = np.zeros((8, 3), np.float32)
    sc = scenes.Scene(name="map_vs_analytic", x=x, v=np.zeros_like(x), boundaries=[analytic, mapped],
                      grid_offset=(-12, -12, -12), grid_res=(24, 24, 24))
    orc = of.Oracle(sc, precision=8)
    dmin, dmax, res, cs = orc.grid_info(1)
    nodes = of.node_positions(dmin, res, cs)
    orc.set_distance_map(1, scenes.shape_distance("box", box, nodes))
    orc.build_volume_map(0)
    orc.build_volume_map(1)
    # points within h of the +x face's centre, inside the container: every
    # quadrature point and map node they reach has the face as its nearest
    # feature, so the distance is linear there
    rng = np.random.default_rng(5)
    pts = np.stack([0.1 - rng.uniform(0.2 * h, 0.9 * h, 50), rng.uniform(-0.02, 0.02, 50),
                    rng.uniform(-0.015, 0.015, 50)], 1)
    rec, ker = orc.sample_boundary(pts)
    np.testing.assert_allclose(rec[1, :, :3], rec[0, :, :3], rtol=0, atol=1e-12)
    assert np.all(rec[0, :, 3] > 0)
    np.testing.assert_allclose(rec[1, :, 3], rec[0, :, 3], rtol=1e-9)
    np.testing.assert_allclose(ker[1], ker[0], rtol=1e-8, atol=1e-9 * np.abs(ker[0]).max())
    # the boundary point is the particle's projection onto the face
    np.testing.assert_allclose(rec[1, :, 0], 0.1, atol=1e-12)
    np.testing.assert_allclose(rec[1, :, 1:3], pts[:, 1:3], atol=1e-12)


def test_map_record_outside_the_map_is_empty():
    box = (0.1, 0.1, 0.1, 0.0)
    body = scenes.BoundarySpec("box", box, 1.0, (8, 8, 8), sampled=True)
    x = np.zeros((4, 3), np.float32)
    sc = scenes.Scene(name="map_outside", x=x, v=np.zeros_like(x), boundaries=[body],
                      grid_offset=(-12, -12, -12), grid_res=(24, 24, 24))
    orc = of.Oracle(sc, precision=8)
    dmin, dmax, res, cs = orc.grid_info(0)
    nodes = of.node_positions(dmin, res, cs)
    orc.set_distance_map(0, scenes.shape_distance("box", box, nodes))
    orc.build_volume_map(0)
    far = np.array([[0.2, 0.0, 0.0], [0.0, -0.09, 0.0]])
    rec, ker = orc.sample_boundary(far)
    np.testing.assert_array_equal(rec[0, :, :3], far)
    assert not rec[0, :, 3].any() and not ker.any()
