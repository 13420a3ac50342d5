"""The numpy access that Depot.coat adds to device variables (get, set,
type_enum_to_dtype), on the CPU: a host-memory stand-in with the bound
classes' byte interface (get_bytes / set_bytes with an element offset,
variable.hpp:60-82 of the reference) takes the place of device memory."""
import numpy as np
import pytest

from alluvion_amd import _alluvion as al
from alluvion_amd import variable


class HostVariable:
    """Host stand-in for a bound Variable class: shape, scalar type,
    components per element and a byte buffer."""

    def __init__(self, other=None, shape=(4,), kind=None, comps=1):
        if other is not None:  # the bound classes' copy constructor
            shape, kind, comps, self.buf = other.shape, other.kind, other.comps, other.buf
        else:
            size = {al.NumericType.f64: 8}.get(kind, 4)
            self.buf = bytearray(int(np.prod(shape)) * comps * size)
        self.shape, self.kind, self.comps = list(shape), kind, comps

    def get_type(self):
        return self.kind

    def get_shape(self):
        return self.shape

    def get_num_primitives_per_element(self):
        return self.comps

    def _elem_bytes(self):
        return len(self.buf) // int(np.prod(self.shape))

    def get_bytes(self, dst):
        dst[:] = np.frombuffer(bytes(self.buf[:dst.size]), np.ubyte)

    def set_bytes(self, src, offset=0):
        o = offset * self._elem_bytes()
        if o + src.size > len(self.buf):
            raise RuntimeError("out of bounds")
        self.buf[o:o + src.size] = src.tobytes()


@pytest.mark.parametrize("kind,dtype", [(al.NumericType.f32, np.float32),
                                        (al.NumericType.f64, np.float64),
                                        (al.NumericType.i32, np.int32),
                                        (al.NumericType.u32, np.uint32)])
def test_dtype_mapping(kind, dtype):
    v = variable.coat(HostVariable(shape=(3,), kind=kind))
    assert v.type_enum_to_dtype() is dtype


def test_vector_elements_get_a_component_axis():
    v = variable.coat(HostVariable(shape=(2, 5), kind=al.NumericType.f32, comps=3))
    src = np.arange(30, dtype=np.float64).reshape(2, 5, 3)
    v.set(src)  # converted to float32
    out = v.get()
    assert out.dtype == np.float32 and out.shape == (2, 5, 3)
    np.testing.assert_array_equal(out, src.astype(np.float32))
    assert v.get(10).shape == (10, 3)  # an explicit shape
    assert v.get([2, 5]).shape == (2, 5, 3)


def test_scalar_elements_and_offsets():
    v = variable.coat(HostVariable(shape=(6,), kind=al.NumericType.u32))
    v.set(np.array([7, 8], np.int64), offset=3)  # element offset
    np.testing.assert_array_equal(v.get(), [0, 0, 0, 7, 8, 0])
    assert v.get().dtype == np.uint32
    with pytest.raises(RuntimeError):
        v.set(np.ones(4), offset=3)


def test_non_contiguous_source_and_coat_idempotence():
    base = HostVariable(shape=(4,), kind=al.NumericType.f32, comps=2)
    v = variable.coat(base)
    src = np.arange(16, dtype=np.float32).reshape(4, 4)[:, ::2]
    v.set(src)
    np.testing.assert_array_equal(v.get(), src)
    assert variable.coat(v) is v
    assert type(v).__name__ == "CoatedHostVariable"
    assert type(variable.coat(HostVariable(shape=(1,), kind=al.NumericType.f32))) is type(v)
    assert not v.is_pinned and not v.is_graphical
