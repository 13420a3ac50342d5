"""Sampling and statistics on the GPU against the oracle (FP32): sample
neighbour lists bit for bit, sample boundary records, the sampled density,
quantities, velocity and vorticity with moving bodies, the 256-bin histogram
bit for bit and the KL divergence."""
import os
import sys

import numpy as np
import pytest

import oracle_ffi as of
from alluvion_amd import scenes

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import histogram_oracle as ho  # noqa: E402

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


@pytest.fixture(scope="module")
def state():
    sc = scenes.dam_break_rigid(nx=32, ny=32, nz=16, body_resolution=16)
    dp, pile, solver, cn, cni = scenes.build_solver(sc)
    for _ in range(30):
        solver.step()
    solver.update_particle_neighbors()
    n = sc.n
    orc = of.Oracle(sc, precision=4)
    for b, var in enumerate(pile.volume_grids):
        orc.set_volume_map(b, dp.coat(var).get())
    for b in range(pile.get_size()):
        q = pile.get_q(b)
        orc.set_rigid(b, list(pile.get_x(b)), list(pile.get_v(b)), [q.x, q.y, q.z, q.w],
                      list(pile.get_omega(b)))
    orc.set("x", dp.coat(solver.particle_x).get()[:n])
    orc.set("v", dp.coat(solver.particle_v).get()[:n])
    orc.compute_all_boundaries()
    orc.update_particle_neighbors()
    rng = np.random.default_rng(8)
    x = dp.coat(solver.particle_x).get()[:n]
    lo, hi = x.min(0) - 0.02, x.max(0) + 0.02
    samples = rng.uniform(lo, hi, (700, 3)).astype(np.float32)
    samples[:50] = x[rng.choice(n, 50, replace=False)]
    return dict(sc=sc, dp=dp, pile=pile, solver=solver, orc=orc, samples=samples, cn=cn)


def test_sample_lists_boundaries_and_fields_match_oracle(state):
    dp, pile, solver, orc, samples = (state[k] for k in ("dp", "pile", "solver", "orc", "samples"))
    n, ns, K, B = state["sc"].n, samples.shape[0], state["sc"].max_num_neighbors, pile.get_size()
    runner = dp.Runner()
    sx = dp.create_coated((ns,), 3)
    sx.set(samples)
    nbr = dp.create((ns, K), 4)
    cnt = dp.create((ns,), 1, np.uint32)
    runner.launch_make_neighbor_list(sx, solver.pid, solver.pid_length, nbr, cnt, solver.grid_anomaly,
                                     ns)
    ids_o, xixj_o, cnt_o = orc.sample_neighbors(samples)
    counts = dp.coat(cnt).get()
    np.testing.assert_array_equal(counts, cnt_o)
    lists = dp.coat(nbr).get().reshape(ns, K, 4)
    for k in range(ns):
        c = counts[k]
        np.testing.assert_array_equal(lists[k, :c, 3].view(np.uint32), ids_o[k, :c], err_msg=str(k))
        np.testing.assert_array_equal(lists[k, :c, :3], xixj_o[k, :c].astype(np.float32),
                                      err_msg=str(k))
    rec = dp.create((B, ns), 4)
    ker = dp.create((B, ns), 4)
    solver.sample_all_boundaries(sx, rec, ker, ns)
    rec_o, ker_o = orc.sample_boundary(samples)
    assert rel(dp.coat(rec).get().reshape(B, ns, 4), rec_o) < 1e-5
    assert rel(dp.coat(ker).get().reshape(B, ns, 4), ker_o) < 1e-5
    ker_g = dp.coat(ker).get().reshape(B, ns, 4).astype(np.float64)
    rec_g = dp.coat(rec).get().reshape(B, ns, 4).astype(np.float64)

    dens = dp.coat(solver.particle_density).get()[:n]
    assert rel(dens, orc.get("density")[:n]) < 1e-5
    out1 = dp.create((ns,), 1)
    runner.launch_sample_fluid_density(nbr, cnt, out1, ns)
    assert rel(dp.coat(out1).get(), orc.sample_density(xixj_o, cnt_o)) < 1e-5
    runner.launch_sample_density(sx, nbr, cnt, out1, ker, ns)
    assert rel(dp.coat(out1).get(), orc.sample_density(xixj_o, cnt_o, ker_g)) < 1e-5
    runner.launch_sample_fluid(sx, solver.particle_x, solver.particle_density, solver.particle_density,
                               nbr, cnt, out1, ns)
    assert rel(dp.coat(out1).get(), orc.sample_fluid(ids_o, xixj_o, cnt_o, dens.astype(np.float64))) < 1e-5
    out3 = dp.create((ns,), 3)
    runner.launch_sample_fluid(sx, solver.particle_x, solver.particle_density, solver.particle_v, nbr,
                               cnt, out3, ns)
    v = dp.coat(solver.particle_v).get()[:n].astype(np.float64)
    assert rel(dp.coat(out3).get(), orc.sample_fluid(ids_o, xixj_o, cnt_o, v)) < 1e-5

    # velocity and vorticity with the bodies' device kinematics of the last step
    rx, rv, rw = (dp.coat(getattr(pile, a)).get().reshape(-1, 3).astype(np.float64)
                  for a in ("x_device", "v_device", "omega_device"))
    assert np.abs(rv).max() > 0 or np.abs(rw).max() > 0  # the bodies move
    sv = dp.create((ns,), 3)
    runner.launch_sample_velocity(sx, solver.particle_x, solver.particle_density, solver.particle_v,
                                  nbr, cnt, sv, rec, ker, pile.x_device, pile.v_device,
                                  pile.omega_device, ns)
    sv_o = orc.sample_velocity(ids_o, xixj_o, cnt_o, rec_g, ker_g, rx, rv, rw)
    assert rel(dp.coat(sv).get(), sv_o) < 1e-5
    vort = dp.create((ns,), 3)
    svg = dp.coat(sv).get().astype(np.float64)
    runner.launch_sample_vorticity(sx, solver.particle_x, solver.particle_density, solver.particle_v,
                                   nbr, cnt, sv, vort, rec, ker, pile.x_device, pile.v_device,
                                   pile.omega_device, ns)
    vort_o = orc.sample_velocity(ids_o, xixj_o, cnt_o, rec_g, ker_g, rx, rv, rw, sample_v=svg)
    assert rel(dp.coat(vort).get(), vort_o) < 1e-4


def test_histogram_and_kl_match_oracle(state):
    dp = state["dp"]
    rng = np.random.default_rng(2)
    v = rng.normal(1.0, 2.0, 100003).astype(np.float32)
    v[:4] = [6.0, -9.0, np.nan, 1e30]
    vv = dp.create_coated((v.size,), 1)
    vv.set(v)
    mask = (rng.uniform(size=v.size) < 0.6).astype(np.float32)
    mask[5] = 2.0  # only exactly 1 counts
    mm = dp.create_coated((v.size,), 1)
    mm.set(mask)
    hist = dp.create((256,), 1, np.uint32)
    hist2 = dp.create((256,), 1, np.uint32)
    runner = dp.Runner()
    runner.launch_histogram256(None, hist, None, vv, -2.0, 6.0, v.size)
    np.testing.assert_array_equal(dp.coat(hist).get(), ho.histogram256(v, -2.0, 6.0))
    runner.launch_histogram256_with_mask(None, hist2, None, mm, vv, -2.0, 6.0, v.size)
    np.testing.assert_array_equal(dp.coat(hist2).get(), ho.histogram256(v, -2.0, 6.0, mask))
    hp, hq = dp.coat(hist).get(), dp.coat(hist2).get()
    kl = runner.calculate_kl_divergence(hist, hist2, int(hp.sum()), int(hq.sum()), 1e-6, 256)
    assert kl == pytest.approx(ho.kl_divergence(hp, hq, int(hp.sum()), int(hq.sum()), 1e-6, 256),
                               rel=1e-5, abs=1e-7)
