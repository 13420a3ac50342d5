"""The .alu variable file format (variable.hpp:132-261, store.cu:47-66) and
the rigid kinematics file (pile.hpp:757-791).

An .alu file is D u32 shape items (outermost first), a u32 0, the u32 number
of primitives per element, a one-byte type label ('f', 'd', 'i', 'u') and
the raw little-endian elements.  The expected bytes are built here with
numpy from that layout; the header reader runs on the CPU, the variable and
pile round trips through device memory need the GPU.
"""
import struct

import numpy as np
import pytest

from alluvion_amd import _alluvion as _al


def alu_bytes(shape, prims, label, payload):
    head = struct.pack("<%dI" % len(shape), *shape) + struct.pack("<II", 0, prims) + label
    return head + np.ascontiguousarray(payload).tobytes()


def test_header_reader(tmp_path):
    f = tmp_path / "x.alu"
    f.write_bytes(alu_bytes((7, 5), 3, b"f", np.zeros((7, 5, 3), np.float32)))
    shape, prims, label = _al.Store.get_alu_info(str(f))
    assert list(shape) == [7, 5] and prims == 3 and label == "f"
    f.write_bytes(struct.pack("<I", 4))
    with pytest.raises(RuntimeError):
        _al.Store.get_alu_info(str(f))
    with pytest.raises(RuntimeError):
        _al.Store.get_alu_info(str(tmp_path / "missing.alu"))


@pytest.mark.gpu
def test_variable_write_matches_format_and_reads_back(tmp_path):
    from alluvion_amd.depot import Depot

    dp = Depot()
    rng = np.random.default_rng(4)
    data = rng.standard_normal((10, 3)).astype(np.float32)
    var = dp.create_coated((10,), 3)
    var.set(data)
    f = tmp_path / "x.alu"
    var.write_file(str(f))
    assert f.read_bytes() == alu_bytes((10,), 3, b"f", data)
    # four rows starting at element 3
    var.write_file(str(f), 4, 3)
    assert f.read_bytes() == alu_bytes((4,), 3, b"f", data[3:7])
    other = dp.create_coated((10,), 3)
    other.set(np.full((10, 3), 9.0, np.float32))
    assert other.read_file(str(f)) == 4
    got = other.get()
    np.testing.assert_array_equal(got[:4], data[3:7])
    assert np.all(got[4:] == 9.0)
    # element offsets in get_bytes / set_bytes
    other.set(data[:2], offset=8)
    np.testing.assert_array_equal(other.get()[8:], data[:2])
    with pytest.raises(Exception):
        var.write_file(str(f), 11)
    with pytest.raises(Exception):
        var.write_file(str(f), 8, 3)


@pytest.mark.gpu
def test_variable_read_converts_fp64_and_rejects_mismatches(tmp_path):
    from alluvion_amd.depot import Depot

    dp = Depot()
    src = np.linspace(-1, 1, 24).reshape(4, 2, 3)
    var = dp.create_coated((4, 2), 3)
    f = tmp_path / "d.alu"
    f.write_bytes(alu_bytes((3, 2), 3, b"d", src[:3]))
    assert var.read_file(str(f)) == 3
    np.testing.assert_array_equal(var.get()[:3], src[:3].astype(np.float32))
    bad = [alu_bytes((5, 2), 3, b"f", np.zeros((5, 2, 3), np.float32)),  # outermost too big
           alu_bytes((2, 3), 3, b"f", np.zeros((2, 3, 3), np.float32)),  # inner shape
           alu_bytes((2,), 3, b"f", np.zeros((2, 3), np.float32)),  # rank
           alu_bytes((2, 2), 4, b"f", np.zeros((2, 2, 4), np.float32)),  # element width
           alu_bytes((2, 2), 3, b"i", np.zeros((2, 2, 3), np.int32)),  # type
           alu_bytes((2, 2), 3, b"f", np.zeros((1, 2, 3), np.float32))]  # truncated
    for k, b in enumerate(bad):
        f.write_bytes(b)
        with pytest.raises(RuntimeError):
            var.read_file(str(f))
            pytest.fail("accepted bad file %d" % k)


@pytest.mark.gpu
def test_pile_kinematics_file_round_trip(tmp_path):
    from alluvion_amd import scenes

    sc = scenes.dam_break_small()
    sc.boundaries[0].resolution = (8, 8, 8)
    sc.boundaries.append(scenes.BoundarySpec("sphere", (0.02,), 1.0, (8, 8, 8), mass=0.1,
                                             inertia=(1e-5, 1e-5, 1e-5), x=(0.0, 0.1, 0.0)))
    dp, pile, solver, _, _ = scenes.build_solver(sc)
    pile.set_x(1, dp.f3(0.1, 0.2, 0.3))
    pile.set_v(1, dp.f3(-1.0, 0.5, 0.25))
    pile.set_q(1, dp.f4(0.0, 0.6, 0.0, 0.8))
    pile.set_omega(1, dp.f3(2.0, 0.0, -1.0))
    f = tmp_path / "pile.bin"
    pile.write_file(str(f), 2.0, 3.0, 4.0)
    raw = np.frombuffer(f.read_bytes(), np.float32).reshape(2, 13)
    np.testing.assert_array_equal(raw[1], np.float32([0.2, 0.4, 0.6, -3.0, 1.5, 0.75,
                                                      0.0, 0.6, 0.0, 0.8, 8.0, 0.0, -4.0]))
    assert _al.Pilefloat.get_size_from_file(str(f)) == 2
    # record 1 of the file into body 0
    assert pile.read_file(str(f), 1, 0, 1) == 1
    np.testing.assert_array_equal(list(pile.get_x(0)), raw[1, 0:3])
    np.testing.assert_array_equal(list(pile.get_v(0)), raw[1, 3:6])
    np.testing.assert_array_equal([pile.get_q(0).x, pile.get_q(0).y, pile.get_q(0).z,
                                   pile.get_q(0).w], raw[1, 6:10])
    np.testing.assert_array_equal(list(pile.get_omega(0)), raw[1, 10:13])
    # reading past the end of the file stops there
    assert pile.read_file(str(f), -1, 0, 1) == 1
