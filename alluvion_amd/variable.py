"""Coated variables: numpy ``get``/``set`` on device variables.

Mirror of the reference's alluvion/variable.py:6-59.
"""
import numpy as np

from . import _alluvion as _al


def type_enum_to_dtype_func(self):
    numeric_type = self.get_type()
    if numeric_type == _al.NumericType.f32:
        return np.float32
    if numeric_type == _al.NumericType.f64:
        return np.float64
    if numeric_type == _al.NumericType.i32:
        return np.int32
    if numeric_type == _al.NumericType.u32:
        return np.uint32
    return numeric_type


def get_func(self, shape=None):
    dtype = self.type_enum_to_dtype()
    if shape is None:
        shape = self.get_shape()
    if not isinstance(shape, (list, tuple)):
        shape = [shape]
    dst = np.empty(int(np.prod(shape)) * self.get_num_primitives_per_element(), dtype)
    self.get_bytes(dst.view(np.ubyte))
    if self.get_num_primitives_per_element() == 1:
        return dst.reshape(shape)
    return dst.reshape(*shape, -1)


def set_func(self, src, offset=0):
    dtype = self.type_enum_to_dtype()
    src = np.ascontiguousarray(src)
    if src.dtype != dtype:
        src = src.astype(dtype)
    self.set_bytes(src.reshape(-1).view(np.ubyte), offset)


variable_class_dict = {}
for _name in dir(_al):
    if not _name.startswith("Variable"):
        continue
    variable_class_dict[_name] = type(
        "Coated" + _name, (getattr(_al, _name),), {
            "type_enum_to_dtype": type_enum_to_dtype_func,
            "get": get_func,
            "set": set_func,
            "is_graphical": False,
            "is_pinned": False,
        })
