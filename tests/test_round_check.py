"""The HIP kernels' correctly rounded square root is built from the
hardware's approximate one (sqrt_rn in alluvion_amd/csrc/alv_device.hpp).  The
hardware result is only known to within one unit in the last place, so this
CPU check runs the same instruction sequence from each of the three floats
nearest the exact root and requires the IEEE result from all three, for
every float in [2^-96, 2^32] and zero.  The GPU's own results are checked in
tests/test_gpu_math.py."""
import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_sqrt_exhaustive(tmp_path):
    out = str(tmp_path / "librc.so")
    subprocess.run(["gcc", "-O2", "-mfma", "-ffp-contract=off", "-fno-fast-math", "-shared",
                    "-fPIC", "-o", out, os.path.join(HERE, "round_check.c"), "-lm"], check=True)
    L = ctypes.CDLL(out)
    fp = ctypes.POINTER(ctypes.c_float)
    L.check_sqrt.argtypes = [ctypes.c_float, ctypes.c_float, fp]
    L.check_sqrt.restype = ctypes.c_longlong
    bad = ctypes.c_float()
    assert L.check_sqrt(2.0 ** -96, 2.0 ** 32, ctypes.byref(bad)) == 0, bad.value
