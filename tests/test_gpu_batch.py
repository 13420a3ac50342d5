"""Batched scenes: one solver holding several independent dam breaks gives,
scene by scene, exactly the results of separate solvers (same kernels, same
per-scene block decomposition and reduction order), including every step's
iteration counts."""
import numpy as np
import pytest

from alluvion_amd import scenes

pytestmark = pytest.mark.gpu

SPECS = [(3000, 0.25), (9000, 0.4), (5000, 0.3), (700, 0.2)]


def make_scenes():
    return [scenes.dam_break_box(n, e, 100 + k, map_res=16) for k, (n, e) in enumerate(SPECS)]


def test_batch_matches_single_scene_solvers():
    scs = make_scenes()
    dp, piles, batch, cn, cni = scenes.build_batch(scs)
    assert batch.num_scenes == len(scs)
    singles = [scenes.build_solver(sc) for sc in make_scenes()]
    for step in range(25):
        batch.step()
        for k, (sdp, spile, ssolver, _, _) in enumerate(singles):
            ssolver.step()
            rep = batch.scene_report(k)
            assert rep["num_divergence_solve"] == ssolver.num_divergence_solve, (step, k)
            assert rep["num_density_solve"] == ssolver.num_density_solve, (step, k)
            assert rep["mean_density_error"] == ssolver.mean_density_error, (step, k)
    X = dp.coat(batch.particle_x).get()
    V = dp.coat(batch.particle_v).get()
    D = dp.coat(batch.particle_density).get()
    for k, (sdp, spile, ssolver, _, _) in enumerate(singles):
        first, count = batch.scene_range(k)
        assert count == scs[k].n
        np.testing.assert_array_equal(X[first:first + count], sdp.coat(ssolver.particle_x).get()[:count])
        np.testing.assert_array_equal(V[first:first + count], sdp.coat(ssolver.particle_v).get()[:count])
        np.testing.assert_array_equal(D[first:first + count],
                                      sdp.coat(ssolver.particle_density).get()[:count])
        assert abs(batch.scene_report(k)["t"] - ssolver.t) < 1e-9


def test_batch_scenes_converge_independently():
    """Scenes need different iteration counts; each stops on its own test."""
    scs = make_scenes()
    for sc in scs:
        scenes.adaptive_step(sc)
    dp, piles, batch, cn, cni = scenes.build_batch(scs)
    counts = set()
    for _ in range(40):
        batch.step()
        reps = [batch.scene_report(k) for k in range(len(scs))]
        for r in reps:
            assert (r["mean_density_error"] <= scs[0].density_error_tolerance or
                    r["num_density_solve"] == scs[0].max_density_solve)
        counts.add(tuple(r["num_density_solve"] for r in reps))
        assert batch.num_density_solve == max(r["num_density_solve"] for r in reps)
    assert any(len(set(c)) > 1 for c in counts)


@pytest.mark.parametrize("slack", [0, 2])
def test_graph_step_matches_direct_launches(slack):
    """The step replayed from captured graphs (30 replays, host reads and
    copies between them) gives exactly the directly launched results."""
    dpa, _, sa, _, _ = scenes.build_solver(scenes.dam_break_small())
    dpb, _, sb, _, _ = scenes.build_solver(scenes.dam_break_small())
    sb.set_graph(True, slack)
    for step in range(30):
        sa.step()
        sb.step()
        assert sa.num_divergence_solve == sb.num_divergence_solve, step
        assert sa.num_density_solve == sb.num_density_solve, step
        if step % 10 == 0:
            np.testing.assert_array_equal(dpa.coat(sa.particle_density).get(),
                                          dpb.coat(sb.particle_density).get())
    np.testing.assert_array_equal(dpa.coat(sa.particle_x).get(), dpb.coat(sb.particle_x).get())
    np.testing.assert_array_equal(dpa.coat(sa.particle_v).get(), dpb.coat(sb.particle_v).get())


def test_graph_batch_matches_direct_batch():
    dpa, _, ba, _, _ = scenes.build_batch(make_scenes())
    dpb, _, bb, _, _ = scenes.build_batch(make_scenes())
    bb.set_graph(True, 1)
    for step in range(20):
        ba.step()
        bb.step()
        for k in range(len(SPECS)):
            assert ba.scene_report(k) == bb.scene_report(k), (step, k)
    np.testing.assert_array_equal(dpa.coat(ba.particle_x).get(), dpb.coat(bb.particle_x).get())


def test_lds_tiles_match_global_gathers():
    """Neighbour passes reading each block's neighbourhood from LDS give the
    results of passes gathering from global memory, bit for bit, for a single
    scene with rigid bodies and for a batch."""
    sc_a = scenes.dam_break_rigid(nx=32, ny=32, nz=16, body_resolution=16)
    sc_b = scenes.dam_break_rigid(nx=32, ny=32, nz=16, body_resolution=16)
    dpa, _, sa, _, _ = scenes.build_solver(sc_a)
    dpb, _, sb, _, _ = scenes.build_solver(sc_b)
    sa.set_tiles(True)
    sb.set_tiles(False)
    for step in range(25):
        sa.step()
        sb.step()
        assert (sa.num_divergence_solve, sa.num_density_solve) == \
            (sb.num_divergence_solve, sb.num_density_solve), step
    for f in ("particle_x", "particle_v", "particle_density", "particle_kappa", "particle_kappa_v"):
        np.testing.assert_array_equal(dpa.coat(getattr(sa, f)).get(),
                                      dpb.coat(getattr(sb, f)).get(), err_msg=f)
    dpa, _, ba, _, _ = scenes.build_batch(make_scenes())
    dpb, _, bb, _, _ = scenes.build_batch(make_scenes())
    ba.set_tiles(True)
    bb.set_tiles(False)
    for step in range(15):
        ba.step()
        bb.step()
    np.testing.assert_array_equal(dpa.coat(ba.particle_x).get(), dpb.coat(bb.particle_x).get())
    np.testing.assert_array_equal(dpa.coat(ba.particle_v).get(), dpb.coat(bb.particle_v).get())


def test_lds_tiles_with_particles_outside_the_grid():
    """Blocks holding particles outside the grid fall back to global
    gathers: their in-grid neighbour cells lie outside the block's window."""
    def scene():
        sc = scenes.dam_break_small()
        sc.grid_offset = (-19, -21, -21)
        sc.grid_res = (38, 42, 42)  # x in [-0.19, 0.19): the wall layers are outside
        return sc
    sc = scene()
    outside = np.sum((sc.x[:, 0] < -0.19) | (sc.x[:, 0] >= 0.19))
    assert outside > 0
    dpa, _, sa, _, _ = scenes.build_solver(scene())
    dpb, _, sb, _, _ = scenes.build_solver(scene())
    sa.set_tiles(True)
    sb.set_tiles(False)
    for step in range(5):
        sa.step()
        sb.step()
    for f in ("particle_x", "particle_v", "particle_density"):
        np.testing.assert_array_equal(dpa.coat(getattr(sa, f)).get(),
                                      dpb.coat(getattr(sb, f)).get(), err_msg=f)
