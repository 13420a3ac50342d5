"""32-node cubic Lagrange volume maps of the oracle (runner.hpp:1131-1694).

Pinned by: the Kronecker property of the shape functions at the nodes that
get_cells / index_to_node_position address; exact reproduction of cubic
polynomials (values and gradients); the reference's field_op grid
(tests/field_op.cu:11-57); closed forms of the volume integral (deep solid,
far fluid, half space by 1-D quadrature); and an independent numpy
evaluation of the field_gen ellipsoid case (tests/field_gen.cu:16-105),
whose distance function is quadratic and therefore interpolated exactly.
"""
import numpy as np
import pytest
from scipy.integrate import quad
from scipy.special import roots_jacobi, roots_legendre

import oracle_ffi as of
from alluvion_amd import scenes


def local_offset(n):
    """Position of node n of a cell in cell units, from the element layout:
    8 corners (x fastest), then pairs on x-, y-, z-directed edges."""
    if n < 8:
        return np.array([n & 1, (n >> 1) & 1, n >> 2], float)
    m = n - 8
    axis, e, s2 = m // 8, (m % 8) // 2, m % 2
    t = (1 + s2) / 3.0
    if axis == 0:
        return np.array([t, e >> 1, e & 1], float)
    if axis == 1:
        return np.array([e & 1, t, e >> 1], float)
    return np.array([e >> 1, e & 1, t], float)


def test_node_layout_and_kronecker_property():
    res = (5, 4, 3)
    dmin = np.array([-1.3, 0.2, 2.0])
    cs = np.array([0.7, 0.4, 1.1])
    seen = set()
    for k in range(res[2]):
        for j in range(res[1]):
            for i in range(res[0]):
                cells = of.cell_nodes(res, (i, j, k))
                for n, l in enumerate(cells):
                    seen.add(int(l))
                    want = dmin + cs * (np.array([i, j, k]) + local_offset(n))
                    np.testing.assert_allclose(of.node_position(dmin, res, cs, int(l)), want,
                                               atol=1e-12)
    assert seen == set(range(of.num_nodes(res)))
    # N_j(xi_n) = delta_jn at the 32 nodes of the element
    for n in range(32):
        N, _ = of.shape_functions(2 * local_offset(n) - 1)
        want = np.zeros(32)
        want[n] = 1
        np.testing.assert_allclose(N, want, atol=1e-14)


MONOMIALS = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (1, 1, 0), (0, 1, 1),
             (1, 0, 1), (0, 0, 2), (3, 0, 0), (0, 3, 0), (0, 0, 3), (2, 1, 0), (1, 1, 1),
             (0, 1, 2), (1, 2, 0)]


def test_cubic_reproduction_values_and_gradients():
    rng = np.random.default_rng(3)
    nodes_xi = np.array([2 * local_offset(n) - 1 for n in range(32)])
    for _ in range(50):
        xi = rng.uniform(-1, 1, 3)
        N, dN = of.shape_functions(xi)
        assert abs(N.sum() - 1) < 1e-13
        for (a, b, c) in MONOMIALS:
            f = nodes_xi[:, 0] ** a * nodes_xi[:, 1] ** b * nodes_xi[:, 2] ** c
            val = xi[0] ** a * xi[1] ** b * xi[2] ** c
            grad = np.array([
                a * xi[0] ** max(a - 1, 0) * xi[1] ** b * xi[2] ** c,
                b * xi[0] ** a * xi[1] ** max(b - 1, 0) * xi[2] ** c,
                c * xi[0] ** a * xi[1] ** b * xi[2] ** max(c - 1, 0),
            ])
            assert abs(N @ f - val) < 1e-12, (a, b, c)
            np.testing.assert_allclose(dN.T @ f, grad, atol=1e-12)


def test_field_op_grid():
    # tests/field_op.cu:19-25: 11^3 cells on (-23.1, 12.23, -31.44)-(45.1, 103.2, 12.4)
    dmin = np.array([-23.1, 12.23, -31.44])
    dmax = np.array([45.1, 103.2, 12.4])
    res = (11, 11, 11)
    cs = (dmax - dmin) / np.array(res)
    nn = of.num_nodes(res)
    assert nn == 12 ** 3 + 2 * 3 * 11 * 12 ** 2
    pos = np.array([of.node_position(dmin, res, cs, l) for l in range(nn)])
    assert np.all(pos >= dmin - 1e-9) and np.all(pos <= dmax + 1e-9)
    # corner nodes form the lattice, x fastest
    lat = np.stack(np.meshgrid(np.arange(12), np.arange(12), np.arange(12), indexing="ij"), -1)
    lat = lat.transpose(2, 1, 0, 3).reshape(-1, 3)
    np.testing.assert_allclose(pos[:12 ** 3], dmin + cs * lat, atol=1e-12)


def _cubic_W(s, h):
    k = 8 / (np.pi * h ** 3)
    q = s / h
    return np.where(q <= 0.5, k * (6 * q ** 3 - 6 * q ** 2 + 1), np.where(q <= 1, 2 * k * (1 - q) ** 3, 0.0))


def half_space_volume(y0, h):
    """Psi at height y0 above a planar solid (solid for y <= 0)."""
    def chi(s):
        return 1.0 if s <= 0 else float(_cubic_W(s, h) / _cubic_W(0.0, h))
    v, _ = quad(lambda z: np.pi * (h * h - z * z) * chi(y0 + z), -h, h,
                points=[-y0] if -h < -y0 < h else None, epsabs=1e-16, limit=200)
    return v


def test_half_space_and_limits_grid_variant():
    h = 1.0
    res = (6, 6, 6)
    dmin = np.array([-3.0, -3.0, -3.0])
    dmax = -dmin
    cs = (dmax - dmin) / np.array(res)
    nn = of.num_nodes(res)
    pos = np.array([of.node_position(dmin, res, cs, l) for l in range(nn)])
    dist = pos[:, 1].copy()  # signed distance to the plane y = 0
    inner = np.where(np.all(np.abs(pos) <= 3.0 - h - 1e-9, axis=1))[0]
    vol = of.volume_from_distance_map(8, h, dmin, dmax, res, cs, 1.0, dist)
    ball = 4 * np.pi / 3 * h ** 3
    for l in inner:
        y0 = pos[l, 1]
        if y0 <= -h:
            assert abs(vol[l] - ball) < 1e-12 * ball
        elif y0 >= 2 * h:
            assert vol[l] == 0.0
        else:
            assert abs(vol[l] - half_space_volume(y0, h)) < 2e-3 * ball, (l, y0)
    # sign = -1 swaps solid and fluid
    vol_neg = of.volume_from_distance_map(8, h, dmin, dmax, res, cs, -1.0, dist)
    for l in inner:
        y0 = pos[l, 1]
        if -2 * h < y0 < h:
            assert abs(vol_neg[l] - half_space_volume(-y0, h)) < 2e-3 * ball


def test_field_gen_ellipsoid_against_numpy():
    # tests/field_gen.cu:16-47: 11^3 cells, r = 1.25, sign = 1, ellipsoid
    r = 1.25
    margin = 4 * r
    dmin = np.array([-30 - margin, -20 - margin, -40 - margin])
    dmax = np.array([30 + margin, 20 + margin, 40 + margin])
    res = (11, 11, 11)
    cs = (dmax - dmin) / np.array(res)
    nn = of.num_nodes(res)
    pos = np.array([of.node_position(dmin, res, cs, l) for l in range(nn)])

    def f(p):
        return p[..., 0] ** 2 / 625 + p[..., 1] ** 2 / 225 + p[..., 2] ** 2 / 1225 - 1.0

    dist = f(pos)
    # the quadratic is reproduced exactly by the element, so a numpy
    # evaluation of the Peirce integral with the analytic function must agree
    x, w = roots_legendre(30)
    xj, wj = roots_jacobi(15, 0.0, 0.5)
    rk, ck = np.sqrt((xj + 1) / 2), wj / (4 * np.sqrt(2))
    th = 2 * np.pi * np.arange(1, 61) / 60
    sphi = np.sqrt(1 - x * x)
    u = np.stack([sphi[None, :] * np.cos(th)[:, None], sphi[None, :] * np.sin(th)[:, None],
                  np.broadcast_to(x[None, :], (60, 30))], -1)  # (60, 30, 3)
    wts = (2 * np.pi / 60) * w[None, :, None] * ck[None, None, :]
    pts = r * rk[None, None, :, None] * u[:, :, None, :]  # (60, 30, 15, 3)
    rng = np.random.default_rng(5)
    pick = rng.choice(nn, 40, replace=False)
    vol = of.volume_from_distance_map(8, r, dmin, dmax, res, cs, 1.0, dist)
    nonzero = 0
    for l in pick:
        d = f(pos[l] + pts)
        inside = np.all((pos[l] + pts >= dmin) & (pos[l] + pts <= dmax), axis=-1)
        chi = np.where(d <= 0, 1.0, _cubic_W(d, r) / _cubic_W(0.0, r))
        chi = np.where(inside, chi, 0.0)
        want = (wts * chi).sum() * r ** 3
        assert abs(vol[l] - want) <= 1e-9 * (4 * np.pi / 3 * r ** 3), l
        nonzero += want > 0
    assert nonzero > 5
    # FP32 oracle within doctest::Approx-style tolerance of FP64
    vol32 = of.volume_from_distance_map(4, r, dmin, dmax, res, cs, 1.0, dist)
    scale = 4 * np.pi / 3 * r ** 3
    np.testing.assert_allclose(vol32, vol, rtol=0, atol=2e-5 * scale)


def test_analytic_box_container_volume():
    # container: 0.4 m box, sign -1 (solid outside), analytic SDF path
    sc = scenes.dam_break_small()
    sc.boundaries[0].resolution = (4, 4, 4)
    orc = of.Oracle(sc, precision=8)
    orc.build_volume_map(0)
    vol = orc.volume_map(0)
    dmin, dmax, res, cs = orc.grid_info(0)
    np.testing.assert_allclose(dmin, -0.21, atol=1e-15)
    h = sc.kernel_radius
    ball = 4 * np.pi / 3 * h ** 3
    pos = np.array([of.node_position(dmin, res, cs, l) for l in range(len(vol))])
    for l in range(len(vol)):
        p = pos[l]
        inside_depth = 0.2 - np.abs(p).max()  # distance into the fluid region
        if inside_depth >= 2 * h:
            assert vol[l] == 0.0
        elif inside_depth <= -h:
            assert abs(vol[l] - ball) < 1e-12 * ball
        # near a face centre (far from edges): half-space profile
        face = np.argmax(np.abs(p))
        others = [a for a in range(3) if a != face]
        if -h < inside_depth < 2 * h and np.all(np.abs(p[others]) < 0.2 - 2 * h):
            assert abs(vol[l] - half_space_volume(inside_depth, h)) < 2e-3 * ball, l
