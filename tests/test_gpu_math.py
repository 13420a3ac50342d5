"""The GPU's arithmetic under the kernel functions, bit for bit against IEEE
single precision computed by numpy (correctly rounded, no fused operations):
the square root over random and edge arguments, the reciprocal over every
significand (three binades) and the general division over random pairs, the
known-divisor division by h and h^2, and the cubic spline W and gradient
scale over every distance of the gradient's dividing branch (1/2 < r/h <= 1)
and a dense sample of (0, h].  The CPU checks of the same sequences are in
tests/test_round_check.py and tests/test_div_known.py."""
import ctypes

import numpy as np
import pytest

import alv_ctypes

pytestmark = pytest.mark.gpu

F32 = np.float32


def run(L, cn, op, a, b=None):
    a = np.ascontiguousarray(a, F32)
    out = np.empty_like(a)
    fp = ctypes.POINTER(ctypes.c_float)
    bp = None if b is None else np.ascontiguousarray(b, F32).ctypes.data_as(fp)
    rc = L.alv_check_math(ctypes.byref(cn), op, a.ctypes.data_as(fp), bp, out.ctypes.data_as(fp),
                          a.size)
    assert rc == 0, L.alv_last_error()
    return out


def assert_bits(got, want, what):
    g = got.view(np.uint32)
    w = np.asarray(want, F32).view(np.uint32)
    bad = np.nonzero(g != w)[0]
    print("%s: %d values, %d differ" % (what, g.size, bad.size), flush=True)
    assert bad.size == 0, (what, got[bad[:5]], np.asarray(want, F32)[bad[:5]])


def gradient_branch(h):
    """Every float r with 1/2 < RN(r/h) <= 1."""
    h = F32(h)
    lo = (F32(0.5) * h).view(np.uint32) - 8
    hi = h.view(np.uint32) + 8
    r = np.arange(lo, hi + 1, dtype=np.uint32).view(F32)
    q = r / h
    return r[(q > F32(0.5)) & (q <= F32(1.0))]


def numpy_grad_scale(r, h, l):
    r = np.asarray(r, F32)
    h, l = F32(h), F32(l)
    rl = np.sqrt(r * r)
    q = rl / h
    s_in = (q * F32(3.0) - F32(2.0)) / h
    with np.errstate(divide="ignore", invalid="ignore"):
        s_out = (F32(1.0) - q) * (q - F32(1.0)) / rl
    scale = np.where(q <= F32(0.5), s_in, np.where(q <= F32(1.0), s_out, F32(0.0)))
    return (l * scale).astype(F32)


def numpy_w(r2, h, h2, k):
    r2 = np.asarray(r2, F32)
    q2 = r2 / F32(h2)
    q = np.sqrt(r2) / F32(h)
    conj = F32(1.0) - q
    inner = (F32(6.0) * q - F32(6.0)) * q2 + F32(1.0)
    outer = F32(2.0) * conj * conj * conj
    res = np.where(q <= F32(0.5), inner, np.where(q <= F32(1.0), outer, F32(0.0)))
    return (res * F32(k)).astype(F32)


def test_sqrt_matches_ieee():
    L = alv_ctypes.library()
    cn = alv_ctypes.const_for(L, 0.01)
    rng = np.random.default_rng(5)
    x = np.concatenate([
        (2.0 ** rng.uniform(-96, 32, 1 << 22)).astype(F32),
        (rng.random(1 << 22) * 1e-4).astype(F32),  # squared distances below h^2
        np.arange(1 << 20, dtype=F32) ** 2,  # perfect squares
        np.array([0.0, 1.0, 4.0, 1e-4], F32),
    ])
    assert_bits(run(L, cn, 0, x), np.sqrt(x), "sqrt")


def test_reciprocal_every_significand():
    L = alv_ctypes.library()
    cn = alv_ctypes.const_for(L, 0.01)
    sig = np.arange(1 << 23, dtype=np.uint32)
    for e in (127, 127 - 7, 127 + 10):  # [1, 2), [2^-7, 2^-6), [2^10, 2^11)
        y = ((np.uint32(e) << np.uint32(23)) | sig).view(F32)
        assert_bits(run(L, cn, 4, y), F32(1.0) / y, "1 / y, exponent %d" % (e - 127))


def test_division_random_pairs():
    L = alv_ctypes.library()
    cn = alv_ctypes.const_for(L, 0.01)
    rng = np.random.default_rng(7)
    n = 1 << 22
    a = (rng.uniform(1.0, 2.0, n) * 2.0 ** rng.integers(-40, 40, n)
         * rng.choice([-1.0, 1.0], n)).astype(F32)
    b = (rng.uniform(1.0, 2.0, n) * 2.0 ** rng.integers(-40, 40, n)).astype(F32)
    assert_bits(run(L, cn, 5, a, b), a / b, "a / b")


@pytest.mark.parametrize("op", [2, 6])
@pytest.mark.parametrize("h", [0.01, 2.0 ** -7])
def test_gradient_scale(h, op):
    L = alv_ctypes.library()
    cn = alv_ctypes.const_for(L, h)
    rs = np.concatenate([gradient_branch(h), (np.arange(1, 1 << 20, dtype=np.float64) /
                                               (1 << 20) * float(h)).astype(F32)])
    assert_bits(run(L, cn, op, rs), numpy_grad_scale(rs, h, cn.cubic_kernel_l),
                "gradient scale h=%g (op %d)" % (h, op))


@pytest.mark.parametrize("h", [0.01, 0.037])
def test_known_divisor_division_and_w(h):
    L = alv_ctypes.library()
    cn = alv_ctypes.const_for(L, h)
    rng = np.random.default_rng(6)
    a = (rng.random(1 << 22) * 4 * h).astype(F32)
    d = np.full_like(a, F32(h))
    assert_bits(run(L, cn, 1, a, d), a / F32(h), "x / h")
    d2 = np.full_like(a, F32(cn.kernel_radius_sqr))
    assert_bits(run(L, cn, 1, a * a, d2), (a * a) / F32(cn.kernel_radius_sqr), "x / h^2")
    r2 = (rng.random(1 << 22) * F32(cn.kernel_radius_sqr) * 1.1).astype(F32)
    assert_bits(run(L, cn, 3, r2), numpy_w(r2, h, cn.kernel_radius_sqr, cn.cubic_kernel_k), "W")
