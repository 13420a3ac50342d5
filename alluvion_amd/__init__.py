"""MI355X-native SPH solver core with kennychufk/alluvion's Python API.

``from alluvion_amd import Depot`` gives the reference's ``Depot`` workflow
(alluvion/depot.py): dtype-suffixed classes, ``create_cn``, ``coat``.  The
compute path is the HIP library ``libalv.so`` (gfx950) behind the C ABI in
``include/alv/alv.h``; there is no CPU fallback.
"""
import os

_HERE = os.path.dirname(os.path.abspath(__file__))

try:
    from . import _alluvion  # noqa: F401
    from ._alluvion import *  # noqa: F401,F403
except ImportError as exc:  # pragma: no cover - depends on the build
    raise ImportError(
        "alluvion_amd native extension is not built (run `python -m alluvion_amd.build`): %s"
        % exc) from exc

from .depot import Depot  # noqa: E402,F401

LIB_PATH = os.path.join(_HERE, "libalv.so")
