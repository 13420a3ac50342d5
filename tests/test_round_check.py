"""The HIP kernels' correctly rounded square root and reciprocal are built
from the hardware's approximate ones (sqrt_rn and rcp_rn in
alluvion_amd/csrc/alv_device.hpp).  The hardware results are only known to
within one unit in the last place, so these CPU checks run the same
instruction sequences from every estimate within one ulp and require the
IEEE result: the square root for every float in [2^-96, 2^32] and zero, the
reciprocal for every significand.  The reciprocal step misses in exactly one
case, which Markstein's theorem predicts (an all-ones significand with the
estimate on the power of two below 1/y); the GPU test
(tests/test_gpu_math.py) shows which estimate the hardware returns there.
The division of the gradient's outer branch, the known-divisor sequence with
the divisor's correctly rounded reciprocal, is checked for every distance of
that branch at several kernel radii."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = {}


def lib(tmp_path_factory):
    if "L" not in _LIB:
        out = str(tmp_path_factory.mktemp("rc") / "librc.so")
        subprocess.run(["gcc", "-O2", "-mfma", "-ffp-contract=off", "-fno-fast-math", "-shared",
                        "-fPIC", "-o", out, os.path.join(HERE, "round_check.c"), "-lm"],
                       check=True)
        L = ctypes.CDLL(out)
        fp = ctypes.POINTER(ctypes.c_float)
        L.check_sqrt.argtypes = [ctypes.c_float, ctypes.c_float, fp]
        L.check_sqrt.restype = ctypes.c_longlong
        L.check_rcp.argtypes = [fp, fp, ctypes.POINTER(ctypes.c_longlong)]
        L.check_rcp.restype = ctypes.c_longlong
        L.check_div_gradient.argtypes = [ctypes.c_float, fp]
        L.check_div_gradient.restype = ctypes.c_longlong
        _LIB["L"] = L
    return _LIB["L"]


def test_sqrt_exhaustive(tmp_path_factory):
    L = lib(tmp_path_factory)
    bad = ctypes.c_float()
    assert L.check_sqrt(2.0 ** -96, 2.0 ** 32, ctypes.byref(bad)) == 0, bad.value


def test_reciprocal_step_every_significand(tmp_path_factory):
    L = lib(tmp_path_factory)
    bad, bad_r0, tried = ctypes.c_float(), ctypes.c_float(), ctypes.c_longlong()
    miss = L.check_rcp(ctypes.byref(bad), ctypes.byref(bad_r0), ctypes.byref(tried))
    # two estimates within one ulp for almost every significand
    assert tried.value > 2 * (1 << 23) - 16, tried.value
    assert miss == 1, miss
    y = np.float32(bad.value)
    assert y.view(np.uint32) == 0x3fffffff  # significand of all ones
    assert bad_r0.value == 0.5  # the power of two below 1/y


@pytest.mark.parametrize("h", [0.01, 2.0 ** -7, 0.025, 0.0036, 0.1, 1.0])
def test_gradient_division_every_distance(tmp_path_factory, h):
    L = lib(tmp_path_factory)
    bad = ctypes.c_float()
    assert L.check_div_gradient(h, ctypes.byref(bad)) == 0, bad.value
