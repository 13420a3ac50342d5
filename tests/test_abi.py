"""The C ABI library loads on a host without a GPU and exports every entry
point that include/alv/alv.h declares; the pybind11 module imports and carries
the reference's Python class names.  No compute call is made."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "alv", "alv.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(alv_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert len(syms) >= 40
    for s in ("alv_solver_create", "alv_solver_step", "alv_solver_update_particle_neighbors",
              "alv_pile_add", "alv_last_error"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from alluvion_amd import LIB_PATH
    if not os.path.exists(LIB_PATH):
        pytest.fail("libalv.so is not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_error_string_without_gpu_call():
    from alluvion_amd import LIB_PATH
    lib = ctypes.CDLL(LIB_PATH)
    lib.alv_last_error.restype = ctypes.c_char_p
    assert isinstance(lib.alv_last_error(), bytes)


def test_python_module_names():
    import alluvion_amd._alluvion as al
    for name in ("Store", "Runnerfloat", "Constfloat", "ConstiN", "Pilefloat", "SolverDffloat",
                 "SolverIifloat", "Variable1Dfloat3", "Variable1Dfloat", "Variable3Dfloat",
                 "float3", "uint3", "int3", "VolumeMethod", "NumericType"):
        assert hasattr(al, name), name
    for name in ("BoxDistancefloat", "BoxShellDistancefloat", "SphereDistancefloat"):
        assert hasattr(al.dg, name), name
