"""Learned warm start for the DFSPH solver: data collection and training.

The network (``_alluvion.WarmStartNet``, 8 -> 64 -> 64 -> 64 -> 2, SiLU,
bfloat16 on the matrix cores) predicts, for each particle, how far its
converged stored stiffnesses (kappa_v, kappa) lie from the reference's warm
starts (half the previous step's, clamped), from eight local features; the
solver adds the prediction to the reference's warm start, so a network that
predicts 0 reproduces it.  Training pairs come from the solver itself: with
the feature log on, every step records the features at its divergence warm
start and, after the step, the converged stiffnesses minus those warm
starts.  Training runs on the GPU with the library's own forward/backward
kernels and Adam.

Typical use::

    X, Y, n = collect_pairs(dp, solver, steps=300)
    net = train(X, Y, n)
    solver.set_warm_start_net(net)      # later steps start from predictions
    ...
    dp.remove(X); dp.remove(Y)

The solver's convergence tests are unchanged: a poor prediction costs
iterations, never accuracy.
"""
from . import _alluvion


def collect_pairs(dp, solver, steps, on_step=None, stride=1):
    """Runs `steps` solver steps and logs the pairs of every `stride`-th one
    (the feature log is on for those steps only); returns device buffers
    (X [pairs][8], Y [pairs][2]) and the number of pairs."""
    n = solver.num_particles
    logged = (steps + stride - 1) // stride
    X = dp.create((logged * n * 8,), 1)
    Y = dp.create((logged * n * 2,), 1)
    k = 0
    try:
        for i in range(steps):
            log = i % stride == 0
            solver.set_feature_log(log)
            solver.step()
            if log:
                solver.export_training_pairs(X, Y, k * n)
                k += 1
            if on_step is not None:
                on_step(i)
    finally:
        solver.set_feature_log(False)
    return X, Y, logged * n


class Snapshot:
    """Device copy of a DFSPH solver's step-to-step state (positions,
    velocities, stored stiffnesses), its pile's body kinematics, time and
    time step, to take one step twice from the same state."""

    def __init__(self, dp, solver, pile):
        self.dp, self.solver, self.pile = dp, solver, pile
        self.fields = ("particle_x", "particle_v", "particle_kappa", "particle_kappa_v")
        self.copies = []
        for name in self.fields:
            var = getattr(solver, name)
            self.copies.append(dp.create(var.get_shape(), var.get_num_primitives_per_element()))

    def save(self):
        for name, copy in zip(self.fields, self.copies):
            copy.set_from(getattr(self.solver, name))
        p = self.pile
        self.bodies = [(p.get_x(b), p.get_v(b), p.get_q(b), p.get_omega(b))
                       for b in range(p.get_size())]
        self.t, self.dt = self.solver.t, self.solver.dt

    def restore(self):
        for name, copy in zip(self.fields, self.copies):
            getattr(self.solver, name).set_from(copy)
        p = self.pile
        for b, (x, v, q, w) in enumerate(self.bodies):
            p.set_x(b, x)
            p.set_v(b, v)
            p.set_q(b, q)
            p.set_omega(b, w)
        self.solver.t, self.solver.dt = self.t, self.dt

    def release(self):
        for copy in self.copies:
            self.dp.remove(copy)
        self.copies = []


def replay_compare(dp, solver, pile, net, steps, on_step=None):
    """Iterations of the learned and the plain warm start from the same
    states.  Each step of a plain-warm-start trajectory is taken twice from
    one snapshot: first with the network, then, after restoring the
    snapshot, with the reference's warm start, from whose result the
    trajectory continues.  Returns per-step arrays (learned iterations,
    plain iterations, learned ms, plain ms); each step ends with the
    solver's synchronisation, so the wall times are per step."""
    import time

    import numpy as np

    snap = Snapshot(dp, solver, pile)
    its = np.zeros((steps, 2), np.int64)
    ms = np.zeros((steps, 2))
    try:
        for i in range(steps):
            snap.save()
            for k, use in enumerate((net, None)):
                if k == 1:
                    snap.restore()
                solver.set_warm_start_net(use)
                t0 = time.perf_counter()
                solver.step()
                ms[i, k] = 1e3 * (time.perf_counter() - t0)
                its[i, k] = solver.num_divergence_solve + solver.num_density_solve
            if on_step is not None:
                on_step(i)
    finally:
        solver.set_warm_start_net(None)
        snap.release()
    return its[:, 0], its[:, 1], ms[:, 0], ms[:, 1]


def train(X, Y, count, seed=1, schedule=((3000, 1e-3), (1000, 2e-4)), batch=65536,
          net=None):
    """Fits the input/output normalisation to the data set and trains with
    Adam on random mini-batches; returns (net, per-step batch losses)."""
    if net is None:
        net = _alluvion.WarmStartNet(seed)
    net.fit_normalization(X, Y, count)
    losses = []
    for k, (steps, lr) in enumerate(schedule):
        losses.extend(net.train(X, Y, count, batch, steps, lr, seed * 1000 + k))
    return net, losses
