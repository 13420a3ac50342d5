"""alv_sum_f32_noalloc (sums into the library's fixed device scratch) gives
the bits of alv_sum_f32 (sums into memory it allocates), on two calls in a
row, the second with a smaller input, and refuses sizes over its maximum."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def lib():
    from alluvion_amd import LIB_PATH
    L = ctypes.CDLL(LIB_PATH)
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    dbl = ctypes.POINTER(ctypes.c_double)
    for name in ("alv_sum_f32", "alv_sum_f32_noalloc"):
        f = getattr(L, name)
        f.argtypes = [vp, sz, dbl]
        f.restype = ctypes.c_int
    L.alv_malloc.argtypes = [ctypes.POINTER(vp), sz]
    L.alv_malloc.restype = ctypes.c_int
    L.alv_free.argtypes = [vp]
    L.alv_free.restype = ctypes.c_int
    L.alv_memcpy_h2d.argtypes = [vp, vp, sz]
    L.alv_memcpy_h2d.restype = ctypes.c_int
    L.alv_last_error.argtypes = []
    L.alv_last_error.restype = ctypes.c_char_p
    return L


def device_array(L, a):
    p = ctypes.c_void_p()
    assert L.alv_malloc(ctypes.byref(p), a.nbytes) == 0, L.alv_last_error()
    assert L.alv_memcpy_h2d(p, a.ctypes.data, a.nbytes) == 0, L.alv_last_error()
    return p


def sums(L, p, n):
    a, b = ctypes.c_double(), ctypes.c_double()
    assert L.alv_sum_f32(p, n, ctypes.byref(a)) == 0, L.alv_last_error()
    assert L.alv_sum_f32_noalloc(p, n, ctypes.byref(b)) == 0, L.alv_last_error()
    return a.value, b.value


def test_noalloc_sum_matches_allocating_sum():
    L = lib()
    rng = np.random.default_rng(3)
    big = rng.standard_normal(3_000_001).astype(np.float32)
    small = (rng.random(1000) * 7.0).astype(np.float32)
    pb, ps = device_array(L, big), device_array(L, small)
    try:
        for arr, p in ((big, pb), (small, ps)):  # second call: smaller input
            ref, got = sums(L, p, arr.size)
            print("n=%d alloc %.17g noalloc %.17g" % (arr.size, ref, got), flush=True)
            assert got == ref
            want = float(np.sum(arr.astype(np.float64)))
            assert abs(got - want) <= 1e-12 * np.abs(arr).astype(np.float64).sum()
        ref, got = sums(L, ps, 0)
        assert got == ref == 0.0
        out = ctypes.c_double()
        rc = L.alv_sum_f32_noalloc(ps, (1 << 36) + 1, ctypes.byref(out))
        assert rc != 0 and b"exceeds the maximum" in L.alv_last_error()
    finally:
        L.alv_free(pb)
        L.alv_free(ps)
