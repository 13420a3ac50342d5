"""Learned warm start for the DFSPH solver: data collection and training.

The network (``_alluvion.WarmStartNet``, 8 -> 64 -> 64 -> 64 -> 2, SiLU,
bfloat16 on the matrix cores) predicts each particle's converged stored
stiffness (kappa_v, kappa) from eight local features.  Training pairs come
from the solver itself: with the feature log on, every step records the
features at its divergence warm start and, after the step, the converged
stiffness it ended with.  Training runs on the GPU with the library's own
forward/backward kernels and Adam.

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


def collect_pairs(dp, solver, steps, on_step=None):
    """Runs `steps` solver steps with the feature log on and returns device
    buffers (X [pairs][8], Y [pairs][2]) and the number of pairs."""
    n = solver.num_particles
    X = dp.create((steps * n * 8,), 1)
    Y = dp.create((steps * n * 2,), 1)
    solver.set_feature_log(True)
    try:
        for i in range(steps):
            solver.step()
            solver.export_training_pairs(X, Y, i * n)
            if on_step is not None:
                on_step(i)
    finally:
        solver.set_feature_log(False)
    return X, Y, steps * n


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
