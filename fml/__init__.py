"""The reference package's import path, ``fml``, for this build.

The reference's callers write ``import fml``, ``from fml.kernels import
get_atomic_kernels_arad`` and ``from fml.math.distance import
get_l2_distance_arad`` (tests/test_arad.py:8-11, examples/example_arad.py:9-12).
This package makes those names resolve, unchanged, to the modules of
``fml_amd``: each submodule is registered under its ``fml.`` name, so
``fml.kernels`` *is* ``fml_amd.kernels`` (one module object, one library
handle, one tuning table).
"""
import importlib as _importlib
import sys as _sys

import fml_amd as _impl
from fml_amd import *  # noqa: F401,F403
from fml_amd import __all__, __version__  # noqa: F401

_SUBMODULES = (
    "arad", "aras", "data", "fml", "ptp", "krr", "batch", "serving", "synth", "tuning",
    "kernels", "kernels.kernels", "kernels.atomic", "kernels.aras_kernels",
    "math", "math.cho_solve", "math.distance",
    "representations",
)

for _name in _SUBMODULES:
    _mod = _importlib.import_module(_impl.__name__ + "." + _name)
    _sys.modules[__name__ + "." + _name] = _mod
    _parent, _, _leaf = (__name__ + "." + _name).rpartition(".")
    setattr(_sys.modules[_parent], _leaf, _mod)

del _name, _mod, _parent, _leaf
