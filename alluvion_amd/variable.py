"""Numpy access to device variables.

``Depot.coat(var)`` returns the variable as an instance of a subclass of its
own bound class that also carries ``get`` (device -> new numpy array),
``set`` (numpy -> device, at an element offset) and ``type_enum_to_dtype``;
these are the calls reference scripts make on coated variables
(alluvion/variable.py:6-59 and depot.py:112-117 of the reference).

Design: one mixin, ``NumpyAccess``, holds the three methods; the coated
subclass of a bound class is made the first time a variable of that class
is coated and cached.  The copies go through the bound class's
``get_bytes`` / ``set_bytes`` (a byte view of the numpy buffer), so no extra
staging buffer is made.
"""
import numpy as np

from . import _alluvion as _al

# NumericType of the bound classes -> numpy scalar type
_NUMPY_TYPE = {
    _al.NumericType.f32: np.float32,
    _al.NumericType.f64: np.float64,
    _al.NumericType.i32: np.int32,
    _al.NumericType.u32: np.uint32,
}


class NumpyAccess:
    """Mixin for bound Variable classes: copies to and from numpy arrays."""

    is_graphical = False
    is_pinned = False

    def type_enum_to_dtype(self):
        kind = self.get_type()
        return _NUMPY_TYPE.get(kind, kind)

    def get(self, shape=None):
        """A new array with the variable's contents.  The shape is the
        variable's (or `shape`), with a trailing axis for the components
        of vector elements (float3 -> [..., 3])."""
        if shape is None:
            shape = self.get_shape()
        dims = list(shape) if isinstance(shape, (list, tuple)) else [int(shape)]
        comps = self.get_num_primitives_per_element()
        out = np.empty(dims + ([comps] if comps > 1 else []), self.type_enum_to_dtype())
        self.get_bytes(out.reshape(-1).view(np.ubyte))
        return out

    def set(self, src, offset=0):
        """Writes `src` (converted to the variable's scalar type) starting
        at element `offset`."""
        data = np.ascontiguousarray(src, dtype=self.type_enum_to_dtype())
        self.set_bytes(data.reshape(-1).view(np.ubyte), offset)


_coated_classes = {}


def coated_class(base):
    """The subclass of bound class `base` that adds NumpyAccess."""
    cls = _coated_classes.get(base)
    if cls is None:
        cls = type("Coated" + base.__name__, (NumpyAccess, base), {})
        _coated_classes[base] = cls
    return cls


def coat(var):
    """`var` with numpy access (itself if it already has it)."""
    if isinstance(var, NumpyAccess):
        return var
    return coated_class(type(var))(var)
