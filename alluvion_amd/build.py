"""Builds the native parts of alluvion_amd in-tree.

* ``libalv.so``: HIP kernels + C ABI (``include/alv/alv.h``) for gfx950.
* ``_alluvion*.so``: pybind11 module exposing the reference's Python names
  over the C ABI (host-only, built with g++).

Usage: ``python -m alluvion_amd.build`` (or ``build()`` from Python).
Object files go to ``alluvion_amd/_build/``.
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I", INCLUDE,
]
HIP_SOURCES = ["alv_step.hip", "alv_capi.hip", "alv_mlp.hip", "alv_contact.hip", "alv_sample.hip"]
LIB = os.path.join(PKG, "libalv.so")


def _run(cmd):
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), res.stdout))
    return res.stdout


def _compile_hip(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_hip():
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(HIP_SOURCES)) as ex:
        objs = list(ex.map(_compile_hip, HIP_SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_python_module():
    import pybind11

    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(PKG, "_alluvion" + suffix)
    tmp = out + ".tmp"
    py_inc = sysconfig.get_paths()["include"]
    cmd = [
        "g++", "-O2", "-std=c++17", "-fPIC", "-shared",
        "-I", pybind11.get_include(), "-I", py_inc, "-I", INCLUDE,
        os.path.join(CSRC, "pyalluvion.cpp"),
        "-L", PKG, "-lalv", "-Wl,-rpath,$ORIGIN",
        "-o", tmp,
    ]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build():
    lib = build_hip()
    mod = build_python_module()
    return lib, mod


if __name__ == "__main__":
    for path in build():
        print(path)
    sys.exit(0)
