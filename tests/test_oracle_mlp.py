"""The CPU oracle of the warm-start network: exact FP64 gradients checked by
central finite differences, closed forms for a network with zero weights, and
the bfloat16-rounded variant (the device kernels' arithmetic) staying close to
the exact one."""
import numpy as np

import oracle_ffi as of

# parameter blocks: W1 b1 W2 b2 W3 b3 W4 b4
SIZES = [64 * 8, 64, 64 * 64, 64, 64 * 64, 64, 2 * 64, 2]
OFFS = np.concatenate([[0], np.cumsum(SIZES)])


def random_params(seed, scale=0.4):
    rng = np.random.default_rng(seed)
    p = np.zeros(of.MLP_PARAMS, np.float32)
    fan_in = [8, 1, 64, 1, 64, 1, 64, 1]
    for b in range(8):
        sc = scale * np.sqrt(3.0 / fan_in[b])
        p[OFFS[b]:OFFS[b + 1]] = rng.uniform(-sc, sc, SIZES[b])
    return p


def batch(seed, n):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((n, 8)).astype(np.float32),
            rng.standard_normal((n, 2)).astype(np.float32))


def test_parameter_count():
    assert of.lib().orc_mlp_num_params() == of.MLP_PARAMS == OFFS[-1]


def test_gradient_matches_finite_differences():
    p = random_params(1)
    x, t = batch(2, 12)
    loss, g = of.mlp_loss_grad(p, x, t, quantize=False)
    rng = np.random.default_rng(3)
    # a few parameters from every block, including biases
    idx = np.concatenate([rng.choice(np.arange(OFFS[b], OFFS[b + 1]), min(6, SIZES[b]),
                                     replace=False) for b in range(8)])
    gmax = np.abs(g).max()
    for k in idx:
        eps = np.float32(1e-3) * max(np.float32(1.0), abs(p[k]))
        pp, pm = p.copy(), p.copy()
        pp[k] = p[k] + eps
        pm[k] = p[k] - eps
        lp, _ = of.mlp_loss_grad(pp, x, t, quantize=False)
        lm, _ = of.mlp_loss_grad(pm, x, t, quantize=False)
        fd = (lp - lm) / (float(pp[k]) - float(pm[k]))
        assert abs(fd - g[k]) <= 1e-5 * gmax + 1e-4 * abs(g[k]), (k, fd, g[k])


def test_zero_weights_closed_form():
    p = np.zeros(of.MLP_PARAMS, np.float32)
    b3 = np.linspace(-1, 1, 64).astype(np.float32)
    b4 = np.array([0.3, -0.2], np.float32)
    p[OFFS[5]:OFFS[6]] = b3
    p[OFFS[7]:OFFS[8]] = b4
    x, t = batch(4, 10)
    y = of.mlp_forward(p, x, quantize=False)
    np.testing.assert_allclose(y, np.broadcast_to(b4.astype(np.float64), y.shape), rtol=0, atol=0)
    loss, g = of.mlp_loss_grad(p, x, t, quantize=False)
    e = b4[None, :].astype(np.float64) - t
    assert abs(loss - (e ** 2).sum(1).mean()) < 1e-12
    dy = 2 * e / len(x)
    np.testing.assert_allclose(g[OFFS[7]:OFFS[8]], dy.sum(0), rtol=1e-12, atol=1e-15)
    a3 = b3 / (1 + np.exp(-b3.astype(np.float64)))
    np.testing.assert_allclose(g[OFFS[6]:OFFS[7]].reshape(2, 64), np.outer(dy.sum(0), a3),
                               rtol=1e-12, atol=1e-15)
    # nothing flows below layer 4 when W4 = 0
    assert np.all(g[:OFFS[6]] == 0)


def test_bf16_rounding_stays_close():
    p = random_params(5)
    x, t = batch(6, 256)
    y0 = of.mlp_forward(p, x, quantize=False)
    y1 = of.mlp_forward(p, x, quantize=True)
    assert np.abs(y1 - y0).max() <= 0.05 * np.abs(y0).max()
    assert np.abs(y1 - y0).max() > 0  # the rounding is really applied
    _, g0 = of.mlp_loss_grad(p, x, t, quantize=False)
    _, g1 = of.mlp_loss_grad(p, x, t, quantize=True)
    assert np.abs(g1 - g0).max() <= 0.05 * np.abs(g0).max()
