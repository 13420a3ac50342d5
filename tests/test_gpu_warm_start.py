"""Learned warm start on the GPU: the matrix-core network against the CPU
oracle (forward values and parameter gradients with the same bfloat16
rounding points), training on the device, and the solver path with the
network switched in."""
import numpy as np
import pytest

import oracle_ffi as of
from alluvion_amd import Depot, scenes

pytestmark = pytest.mark.gpu

SIZES = [64 * 8, 64, 64 * 64, 64, 64 * 64, 64, 2 * 64, 2]
OFFS = np.concatenate([[0], np.cumsum(SIZES)])
IDENTITY_NORM = np.array([0] * 8 + [1] * 8 + [0, 0, 1, 1], np.float32)


def random_params(seed, scale=0.6):
    rng = np.random.default_rng(seed)
    p = np.zeros(of.MLP_PARAMS, np.float32)
    fan_in = [8, 1, 64, 1, 64, 1, 64, 1]
    for b in range(8):
        sc = scale * np.sqrt(3.0 / fan_in[b])
        p[OFFS[b]:OFFS[b + 1]] = rng.uniform(-sc, sc, SIZES[b])
    return p


def device_array(dp, a):
    a = np.ascontiguousarray(a, np.float32).ravel()
    v = dp.create((a.size,), 1)
    dp.coat(v).set(a)
    return v


def test_forward_matches_oracle():
    import alluvion_amd._alluvion as al
    dp = Depot(np.float32)
    rng = np.random.default_rng(11)
    n = 5000  # not a multiple of the 16-sample tile
    x = rng.standard_normal((n, 8)).astype(np.float32)
    p = random_params(12)
    net = al.WarmStartNet(0)
    net.set_params(p)
    net.set_normalization(IDENTITY_NORM)
    X = device_array(dp, x)
    Y = dp.create((n * 2,), 1)
    net.forward(X, Y, n)
    y = dp.coat(Y).get().reshape(n, 2)
    ref = of.mlp_forward(p, x, quantize=True)
    scale = np.abs(ref).max()
    err = np.abs(y - ref)
    # FP32 accumulation against FP64: differences stay tiny except where a
    # bf16 activation rounds the other way, which moves an output by about
    # one bf16 step of one hidden unit
    assert np.median(err) <= 1e-4 * scale, np.median(err) / scale
    assert err.max() <= 2e-2 * scale, err.max() / scale
    # normalisation: shift/scale on the inputs and the outputs
    norm = np.array(list(np.full(8, 0.5)) + list(np.full(8, 2.0)) + [1.0, -3.0, 10.0, 0.1],
                    np.float32)
    net.set_normalization(norm)
    net.forward(X, Y, n)
    y2 = dp.coat(Y).get().reshape(n, 2)
    ref2 = of.mlp_forward(p, (x - 0.5) * 2.0, quantize=True) * np.array([10.0, 0.1]) + \
        np.array([1.0, -3.0])
    assert np.median(np.abs(y2 - ref2) / np.abs(ref2).max(0)) <= 1e-4


def test_gradient_matches_oracle():
    import alluvion_amd._alluvion as al
    dp = Depot(np.float32)
    rng = np.random.default_rng(21)
    n = 4096
    x = rng.standard_normal((n, 8)).astype(np.float32)
    t = rng.standard_normal((n, 2)).astype(np.float32)
    p = random_params(22)
    net = al.WarmStartNet(0)
    net.set_params(p)
    loss, g = net.loss_grad(device_array(dp, x), device_array(dp, t), n)
    oloss, og = of.mlp_loss_grad(p, x, t, quantize=True)
    assert abs(loss - oloss) <= 1e-4 * oloss, (loss, oloss)
    for b in range(8):
        gb, ob = g[OFFS[b]:OFFS[b + 1]], og[OFFS[b]:OFFS[b + 1]]
        rel = np.linalg.norm(gb - ob) / max(np.linalg.norm(ob), 1e-30)
        assert rel <= 2e-3, (b, rel)
    # the exact-arithmetic gradient differs by the bf16 rounding, much more
    _, eg = of.mlp_loss_grad(p, x, t, quantize=False)
    assert np.linalg.norm(g - eg) > np.linalg.norm(g - og)


def test_training_fits_a_smooth_map():
    import alluvion_amd._alluvion as al
    dp = Depot(np.float32)
    rng = np.random.default_rng(31)
    n = 1 << 18
    x = rng.standard_normal((n, 8)).astype(np.float32)
    t = np.stack([np.sin(x[:, 0]) + 0.5 * x[:, 1], x[:, 2] * x[:, 3] - 0.3 * x[:, 7]], 1)
    t = (3.0 * t + 5.0).astype(np.float32)  # the normalisation must undo this
    X, T = device_array(dp, x), device_array(dp, t)
    net = al.WarmStartNet(3)
    net.fit_normalization(X, T, n)
    loss = net.train(X, T, n, 8192, 400, 3e-3, 5)
    assert np.all(np.isfinite(loss))
    assert loss[-20:].mean() < 0.1 * loss[:5].mean(), (loss[:5], loss[-20:])
    Y = dp.create((n * 2,), 1)
    net.forward(X, Y, n)
    y = dp.coat(Y).get().reshape(n, 2)
    rmse = np.sqrt(((y - t) ** 2).mean(0))
    assert np.all(rmse < 0.35 * t.std(0)), rmse


def test_feature_log_and_learned_step():
    """Features are recorded with their meaning and the targets are the
    converged stiffnesses minus the reference's warm starts; with an all-zero
    network (every correction 0) a step matches the oracle's step with the
    reference's warm starts."""
    import alluvion_amd._alluvion as al
    sc = scenes.dam_break_small()
    dp, pile, solver, cn, cni = scenes.build_solver(sc, "df")
    n = sc.n
    for _ in range(30):
        solver.step()
    solver.set_feature_log(True)
    kv_prev = dp.coat(solver.particle_kappa_v).get()[:n].copy()
    k_prev = dp.coat(solver.particle_kappa).get()[:n].copy()
    v_prev = dp.coat(solver.particle_v).get()[:n].copy()
    solver.step()
    X = dp.create((n * 8,), 1)
    Y = dp.create((n * 2,), 1)
    solver.export_training_pairs(X, Y)
    f = dp.coat(X).get().reshape(n, 8)
    y = dp.coat(Y).get().reshape(n, 2)
    # the pairs come in an internal order: match them by the previous
    # stiffnesses (features 4, 5) to the converged ones minus the reference's
    # warm starts
    kv_now = dp.coat(solver.particle_kappa_v).get()[:n]
    k_now = dp.coat(solver.particle_kappa).get()[:n]
    f32 = np.float32
    want = sorted(zip(kv_prev, k_prev, kv_now - f32(0.5) * np.maximum(kv_prev, f32(-0.5)),
                      k_now - f32(0.5) * np.maximum(k_prev, f32(-0.00025))))
    got = sorted(zip(f[:, 4], f[:, 5], y[:, 0], y[:, 1]))
    assert got == want
    assert np.allclose(np.sort(f[:, 4]), np.sort(kv_prev))
    assert np.allclose(np.sort(f[:, 5]), np.sort(k_prev))
    assert np.allclose(np.sort(f[:, 6]), np.sort(np.linalg.norm(v_prev, axis=1)), rtol=1e-6)
    assert np.all(f[:, 7] == np.float32(sc.dt))
    assert np.all(f[:, 2] == np.floor(f[:, 2])) and f[:, 2].max() <= sc.max_num_neighbors
    assert np.all(f[:, 1] >= 0)
    # an all-zero network predicts no correction: the reference's warm starts
    net = al.WarmStartNet(0)
    net.set_params(np.zeros(of.MLP_PARAMS, np.float32))
    solver.set_warm_start_net(net)
    orc = of.Oracle(sc, precision=4)
    for b, var in enumerate(pile.volume_grids):
        orc.set_volume_map(b, dp.coat(var).get())
    orc.set("x", dp.coat(solver.particle_x).get()[:n])
    orc.set("v", dp.coat(solver.particle_v).get()[:n])
    orc.set("kappa", dp.coat(solver.particle_kappa).get()[:n])
    orc.set("kappa_v", dp.coat(solver.particle_kappa_v).get()[:n])
    solver.step()
    orc.step()
    rep = orc.report()
    assert solver.num_divergence_solve == rep["num_divergence_solve"]
    assert solver.num_density_solve == rep["num_density_solve"]
    for name, var in (("v", solver.particle_v), ("density", solver.particle_density)):
        g = dp.coat(var).get()[:n].astype(np.float64)
        o = orc.get(name)[:n].astype(np.float64)
        assert np.abs(g - o).max() <= 1e-5 * np.abs(o).max(), name
    solver.set_warm_start_net(None)


def test_learned_warm_start_cuts_iterations():
    """Train on 300 plain steps of the 16,384-particle dam break with the
    CFL-limited time step, then take each of 200 further steps twice from the
    same state (replay): with the network and with the reference's warm
    start.  The network needs fewer Jacobi iterations, every learned step
    still meets the unchanged 0.1 % tolerance (or stops at the cap), and a
    replayed plain step repeats the trajectory's own step bit for bit."""
    from alluvion_amd import warm_start

    sc = scenes.adaptive_step(scenes.dam_break_small())
    dp, pile, solver, cn, cni = scenes.build_solver(sc, "df")
    X, Y, count = warm_start.collect_pairs(dp, solver, 300)
    net, losses = warm_start.train(X, Y, count)
    assert np.mean(losses[-100:]) < 0.5 * np.mean(losses[:10])
    dp.remove(X)
    dp.remove(Y)
    # the snapshot restores the state exactly: a plain step taken twice
    snap = warm_start.Snapshot(dp, solver, pile)
    snap.save()
    solver.step()
    x1 = dp.coat(solver.particle_x).get().copy()
    snap.restore()
    solver.step()
    np.testing.assert_array_equal(dp.coat(solver.particle_x).get(), x1)
    snap.release()
    lit, pit, _, _ = warm_start.replay_compare(dp, solver, pile, net, 200)
    floor = sc.min_divergence_solve + sc.min_density_solve  # no step takes fewer
    print("replay iterations per step: plain %.3f learned %.3f (ratio %.3f, floor %d)" % (
        pit.mean(), lit.mean(), pit.mean() / lit.mean(), floor))
    assert pit.mean() > floor + 0.3  # the plain warm start leaves room
    # the network closes at least 3/4 of the plain warm start's distance to
    # the floor of the minimum iteration counts, and is never worse on average
    assert lit.mean() - floor <= 0.25 * (pit.mean() - floor)
    # learned steps converge: re-run a few with the network and check
    solver.set_warm_start_net(net)
    for _ in range(20):
        solver.step()
        assert (solver.mean_density_error <= sc.density_error_tolerance or
                solver.num_density_solve == sc.max_density_solve)
        assert (solver.mean_density_change <= sc.density_change_tolerance / solver.dt or
                solver.num_divergence_solve == sc.max_divergence_solve)
    solver.set_warm_start_net(None)
