"""The HIP kernels' correctly rounded square root and gradient division are
built from the hardware's approximate ones (sqrt_rn, div_rn in
alluvion_amd/csrc/alv_device.hpp).  The hardware results are only known to
within one unit in the last place, so these CPU checks run the same
instruction sequences from each of the three floats nearest the exact value
and require the IEEE result from all three:

* the square root, for every float in [2^-96, 2^32];
* the gradient's (1 - q)(q - 1) / r, for every r of its branch
  (1/2 < RN(r/h) <= 1) at the radii of the workloads, where the library's
  own host check (alv_kernel_division_exact, which decides whether the kernels
  use the fast division) must agree: the sequence is exact for h = 0.01 but
  not for h = 2^-7, whose branch holds a divisor one unit below a power of
  two where a one-unit reciprocal estimate can miss RN(1/r) by more than half
  a unit;
* random quotients, where the sequence is exact too.
The GPU's own results are checked in tests/test_gpu_math.py."""
import ctypes
import os
import subprocess

import pytest

import alv_ctypes

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def chk(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("rc") / "librc.so")
    subprocess.run(["gcc", "-O2", "-mfma", "-ffp-contract=off", "-fno-fast-math", "-shared",
                    "-fPIC", "-o", out, os.path.join(HERE, "round_check.c"), "-lm"], check=True)
    L = ctypes.CDLL(out)
    fp = ctypes.POINTER(ctypes.c_float)
    L.check_sqrt.argtypes = [ctypes.c_float, ctypes.c_float, fp]
    L.check_sqrt.restype = ctypes.c_longlong
    L.check_div_gradient.argtypes = [ctypes.c_float, fp]
    L.check_div_gradient.restype = ctypes.c_longlong
    L.check_div_random.argtypes = [ctypes.c_uint64, ctypes.c_longlong, fp, fp]
    L.check_div_random.restype = ctypes.c_longlong
    return L


def test_sqrt_exhaustive(chk):
    bad = ctypes.c_float()
    assert chk.check_sqrt(2.0 ** -96, 2.0 ** 32, ctypes.byref(bad)) == 0, bad.value


@pytest.mark.parametrize("h,exact", [(0.01, True), (0.02, True), (0.037, True),
                                     (2.0 ** -7, False)])
def test_gradient_division_domain(chk, h, exact):
    bad = ctypes.c_float()
    miss = chk.check_div_gradient(h, ctypes.byref(bad))
    assert (miss == 0) == exact, (miss, bad.value)
    L = alv_ctypes.library()
    cn = alv_ctypes.const_for(L, h)
    flag = ctypes.c_int(-1)
    assert L.alv_kernel_division_exact(ctypes.byref(cn), ctypes.byref(flag)) == 0
    assert flag.value == (1 if exact else 0)


def test_division_random(chk):
    a, b = ctypes.c_float(), ctypes.c_float()
    assert chk.check_div_random(12345, 20_000_000, ctypes.byref(a), ctypes.byref(b)) == 0, \
        (a.value, b.value)
