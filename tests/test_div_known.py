"""The kernels divide by the kernel radius h and by h^2 through a known-
divisor division (Markstein's correction step from the correctly rounded
reciprocal, div_known in alluvion_amd/csrc/alv_device.hpp).  Its results must
be the bits of the IEEE division, which the oracle and the reference use.
Checked here on the CPU with the same three operations (one product, two fused
multiply-adds): exhaustively over every float in the kernels' range for the
radii of the workloads, and on random divisors and dividends."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("divk") / "libdivk.so")
    subprocess.run(["gcc", "-O2", "-mfma", "-ffp-contract=off", "-fno-fast-math", "-shared",
                    "-fPIC", "-o", out, os.path.join(HERE, "div_known_check.c"), "-lm"],
                   check=True)
    L = ctypes.CDLL(out)
    fp = ctypes.POINTER(ctypes.c_float)
    L.check_range.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, fp]
    L.check_range.restype = ctypes.c_longlong
    L.check_random.argtypes = [ctypes.c_uint64, ctypes.c_longlong, fp, fp]
    L.check_random.restype = ctypes.c_longlong
    return L


@pytest.mark.parametrize("h", [0.01, 0.02, 0.0025 * 4, 1.0 / 128, 0.037])
def test_division_by_radius_is_ieee_exhaustive(lib, h):
    """Every float of the quotients' range: distances and squared distances
    up to 2h and 4h^2 down to 2^-24 of them (the kernel arguments r/h, r^2/h^2
    and (3q - 2)/h), both signs."""
    h = np.float32(h)
    h2 = np.float32(h * h)  # kernel_radius_sqr as the constant setters compute it
    bad = ctypes.c_float()
    for d, hi in ((h, 2 * h), (h2, 4 * h2)):
        lo = np.float32(hi) * np.float32(2.0 ** -24)
        miss = lib.check_range(float(d), float(lo), float(hi), ctypes.byref(bad))
        assert miss == 0, (float(d), bad.value)


def test_division_by_known_divisor_random(lib):
    bd, bx = ctypes.c_float(), ctypes.c_float()
    miss = lib.check_random(0x9E3779B97F4A7C15, 20_000_000, ctypes.byref(bd), ctypes.byref(bx))
    assert miss == 0, (bd.value, bx.value)
