"""Pins the oracle's sampling functions (make_neighbor_list on sample points,
runner.hpp:1869-1911; sample_fluid, sample_fluid_density,
sample_position_density, compute_sample_velocity / vorticity,
runner.hpp:2100-2120, 4173-4324) and the histogram / KL restatement
(runner.hpp:746-762, 4425-4539, 4803-4815):

* sample neighbour lists equal a brute-force minimum-image search, in the
  reference order (27 cells, dx outer, dz inner; ascending id inside a
  cell), and skip the particle whose id equals the sample index;
* on a perfect periodic lattice at rest, at lattice sites: the sampled
  density is the closed lattice sum m (W(0) + sum W), the interpolant of a
  constant is that constant and of a linear field is the field (lattice
  symmetry), a uniform velocity is reproduced, and the vorticity of the rigid
  rotation v = w x x is 2 c w with c = sum_j V (grad W)_x d_x over the
  lattice offsets;
* sample boundary records at particle positions equal the particle records,
  and the position density there equals the particle density;
* histograms of values at known bins, the wrap and clamp semantics of the
  quantiser, masks; KL(p, p) = 0 and KL against scipy's relative entropy.
"""
import os
import sys

import numpy as np
import pytest

import oracle_ffi as of
from alluvion_amd import scenes
from test_oracle_solver import R, H, lattice_sum, periodic_lattice

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import histogram_oracle as ho  # noqa: E402


def lattice_oracle(m=20, mass=None):
    S = lattice_sum(2 * R, H)
    mass = 1000.0 / S if mass is None else mass
    sc, x, v = periodic_lattice(m=m, mass=mass)
    orc = of.Oracle(sc, precision=8)
    orc.set("x", x)
    orc.set("v", v)
    orc.compute_all_boundaries()
    orc.update_particle_neighbors()
    return sc, orc, x, mass, S


def interior(x, L, margin):
    return np.nonzero(np.all(np.abs(x) < 0.5 * L - margin, axis=1))[0]


def test_sample_neighbors_match_brute_force_in_reference_order():
    sc, orc, x, _, _ = lattice_oracle()
    L = sc.wrap_length[0]
    rng = np.random.default_rng(5)
    samples = rng.uniform(-0.5 * L, 0.5 * L, (200, 3))
    samples[:20] = x[20:40]  # lattice sites (sample k sits on particle 20 + k)
    samples[40] = x[40]  # sample 40 sits on particle 40: it is skipped
    ids, xixj, counts = orc.sample_neighbors(samples)
    gx = orc.get("x")[: sc.n]
    for k in range(samples.shape[0]):
        d = samples[k] - gx
        d -= L * np.round(d / L)  # minimum image
        r2 = (d ** 2).sum(1)
        want = set(np.nonzero(r2 < H * H)[0]) - {k}
        got = ids[k, :counts[k]]
        assert set(got.tolist()) == want, k
        # reference order: cell offset (dx, dy, dz) row-major, then id
        cs = np.floor(samples[k] / H).astype(int)
        cj = np.floor(gx[got] / H).astype(int)
        off = (cj - cs + 1) % (int(round(L / H)))
        key = (off[:, 0] * 9 + off[:, 1] * 3 + off[:, 2]) * 100000 + got
        assert np.all(np.diff(key) > 0), k
        np.testing.assert_allclose(xixj[k, :counts[k]], (samples[k] - gx[got] + 0.5 * L) % L - 0.5 * L,
                                   atol=1e-12)
    assert 40 not in ids[40, :counts[40]]


def test_lattice_site_samples_reproduce_closed_forms():
    sc, orc, x, mass, S = lattice_oracle()
    L = sc.wrap_length[0]
    idx = interior(x, L, 2 * H)[:60]
    samples = x[idx]
    assert np.all(idx != np.arange(idx.size))  # no sample skips its own site
    ids, xixj, counts = orc.sample_neighbors(samples)
    rho = mass * S
    np.testing.assert_allclose(orc.sample_density(xixj, counts), rho, rtol=1e-12)
    np.testing.assert_allclose(orc.sample_fluid(ids, xixj, counts, np.full(sc.n, 3.5)), 3.5, rtol=1e-12)
    gx = orc.get("x")[: sc.n]
    A = np.array([[0.3, -1.2, 0.7], [2.0, 0.1, -0.4], [0.5, 0.9, 1.1]])
    lin = gx @ A.T + np.array([0.2, -0.1, 0.05])
    np.testing.assert_allclose(orc.sample_fluid(ids, xixj, counts, lin),
                               samples @ A.T + np.array([0.2, -0.1, 0.05]), rtol=0, atol=1e-12)
    # uniform velocity and rigid rotation, no boundaries
    nb0 = np.zeros((0, samples.shape[0], 4))
    z3 = np.zeros((0, 3))
    U = np.array([0.4, -0.2, 0.9])
    orc.set("v", np.tile(U, (sc.n, 1)))
    sv = orc.sample_velocity(ids, xixj, counts, nb0, nb0, z3, z3, z3)
    np.testing.assert_allclose(sv, np.tile(U, (samples.shape[0], 1)), rtol=1e-12)
    w = np.array([0.5, -1.5, 2.0])
    orc.set("v", np.cross(w, gx))
    sv = orc.sample_velocity(ids, xixj, counts, nb0, nb0, z3, z3, z3)
    np.testing.assert_allclose(sv, np.cross(w, samples), rtol=0, atol=1e-12)
    vort = orc.sample_velocity(ids, xixj, counts, nb0, nb0, z3, z3, z3, sample_v=sv)
    # lattice constant c = sum_j V (grad W(d))_x d_x over the offsets d
    g = np.arange(-2, 3) * 2 * R
    D = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    gw, _ = of.cubic_kernel(D, H)
    c = (mass / rho) * np.sum(gw[:, 0] * D[:, 0])
    np.testing.assert_allclose(vort, np.tile(2 * c * w, (samples.shape[0], 1)), rtol=1e-10)


def test_boundary_samples_at_particles_equal_the_particle_records():
    sc = scenes.dam_break_small()
    sc.boundaries[0].resolution = (8, 8, 8)
    orc = of.Oracle(sc, precision=8)
    orc.build_volume_map(0)
    orc.compute_all_boundaries()
    orc.update_particle_neighbors()
    n = sc.n
    pick = np.arange(0, n, 97)
    x = orc.get("x")[:n]
    rec, ker = orc.sample_boundary(x[pick])
    np.testing.assert_array_equal(rec[0], orc.get("boundary").reshape(1, -1, 4)[0, pick])
    np.testing.assert_array_equal(ker[0], orc.get("boundary_kernel").reshape(1, -1, 4)[0, pick])
    # sample indices far from the sampled particles: the sample neighbourhood
    # is the particle's own (itself included, at distance 0)
    samples = x[pick]
    ids, xixj, counts = orc.sample_neighbors(samples)
    far = [k for k in range(len(pick)) if np.linalg.norm(x[k] - samples[k]) > 2 * H]
    dens = orc.sample_density(xixj, counts, ker)
    np.testing.assert_allclose(dens[far], orc.get("density")[pick][far], rtol=1e-12)
    # sample 0 sits on particle 0 and skips it
    assert pick[0] == 0 and 0 not in ids[0, :counts[0]]
    assert dens[0] < orc.get("density")[0]


def test_histogram_bins_wrap_clamp_and_mask():
    lo, hi = -2.0, 6.0
    step = (hi - lo) / 256
    centers = lo + (np.arange(256) + 0.5) * step
    reps = np.arange(256) % 7 + 1
    v = np.repeat(centers, reps).astype(np.float32)
    np.testing.assert_array_equal(ho.histogram256(v, lo, hi), reps)
    # v_max wraps to bin 0, values below v_min and NaN clamp to bin 0, and
    # 257.5 steps above v_min wraps to bin 1
    special = np.array([hi, lo - 1.0, np.nan, lo + 257.5 * step], np.float32)
    h = ho.histogram256(special, lo, hi)
    assert h[0] == 3 and h[1] == 1 and h.sum() == 4
    mask = (np.arange(v.size) % 2).astype(np.float32)
    want = np.bincount(np.repeat(np.arange(256), reps)[mask == 1], minlength=256)
    np.testing.assert_array_equal(ho.histogram256(v, lo, hi, mask), want)


def test_kl_divergence():
    scipy_special = pytest.importorskip("scipy.special")
    rng = np.random.default_rng(3)
    p = rng.integers(0, 50, 256).astype(np.uint32)
    q = rng.integers(1, 50, 256).astype(np.uint32)
    n_p, n_q = int(p.sum()), int(q.sum())
    assert ho.kl_divergence(p, p, n_p, n_p, 1e-6, 256) == pytest.approx(0.0, abs=1e-6)
    ref = float(np.sum(scipy_special.rel_entr(p / n_p, q / n_q)))
    assert ho.kl_divergence(p, q, n_p, n_q, 0.0, 256) == pytest.approx(ref, rel=1e-5)
    # the lower bound stands in for empty q bins
    q0 = q.copy()
    q0[:10] = 0
    got = ho.kl_divergence(p, q0, n_p, n_q, 0.5, 256)
    ref = float(np.sum(scipy_special.rel_entr(p / n_p, np.maximum(q0, 0.5) / n_q)))
    assert got == pytest.approx(ref, rel=1e-5)
