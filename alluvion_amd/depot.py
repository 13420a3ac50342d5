"""Depot: the reference's entry point (alluvion/depot.py:8-117).

Binds dtype-suffixed classes (``Runner``, ``Pile``, ``SolverDf``,
``SolverIi``, ``dg`` distances), the constants helpers and the vector types,
and dispatches ``create(shape, n, dtype)`` to ``create{D}D{dtype}{n}``.
The device path computes in FP32, so only ``np.float32`` is accepted.
"""
from typing import Iterable, Union

import numpy as np

from . import _alluvion as _al
from ._alluvion import Store
from .variable import coat as coat_variable


class Depot(Store):

    def __init__(self, default_dtype=np.float32):
        if default_dtype != np.float32:
            raise NotImplementedError(
                "the MI355X solver core computes in FP32; use np.float32")
        self.default_dtype = default_dtype
        super().__init__()
        suffix = self.dtype_to_string(default_dtype)
        for classname in ["Runner", "Pile", "SolverDf", "SolverIi", "SolverI"]:
            setattr(self, classname, getattr(_al, f"{classname}{suffix}"))
        for classname in ["Distance", "BoxDistance", "BoxShellDistance", "SphereDistance",
                          "MapDistance"]:
            setattr(self, classname, getattr(_al.dg, f"{classname}{suffix}"))
        self.cn = getattr(self, f"cn{suffix}")
        self.create_cn = getattr(self, f"create_cn{suffix}")
        self.copy_cn = getattr(self, f"copy_cn{suffix}")
        self.copy_cn_external = getattr(self, f"copy_cn_external{suffix}")
        for name in ["2", "3", "4"]:
            setattr(self, f"f{name}", getattr(_al, f"{suffix}{name}"))

    def _create(self, shape: Union[int, Iterable[int]], n: int, dtype=None):
        if isinstance(shape, int):
            shape = [shape]
        shape = list(shape)
        postfix = self.get_variable_postfix(len(shape), dtype, n)
        return getattr(super(), f"create{postfix}")(shape)

    def create(self, shape: Union[int, Iterable[int]], n: int, dtype=None):
        return self._create(shape, n, dtype)

    def create_coated(self, shape, n: int, dtype=None):
        return self.coat(self.create(shape, n, dtype))

    def create_coated_like(self, var):
        return self.create_coated(var.get_shape(), var.get_num_primitives_per_element(),
                                  self.coat(var).type_enum_to_dtype())

    def remove(self, var):
        coated = self.coat(var)
        postfix = self.get_variable_postfix(
            dim=len(var.get_shape()), dtype=coated.type_enum_to_dtype(),
            n=var.get_num_primitives_per_element())
        getattr(super(), f"remove{postfix}")(var)

    def dtype_to_string(self, dtype):
        if dtype is None:
            dtype = self.default_dtype
        if dtype == np.float32:
            return "float"
        if dtype == np.float64:
            return "double"
        if dtype == np.uint32:
            return "uint"
        if dtype == np.int32:
            return "int"
        raise NotImplementedError(f"{dtype} not yet implemented.")

    def get_variable_postfix(self, dim, dtype, n):
        return f"{dim}D{self.dtype_to_string(dtype)}{n if n > 1 else ''}"

    @classmethod
    def coat(cls, var):
        return coat_variable(var)
