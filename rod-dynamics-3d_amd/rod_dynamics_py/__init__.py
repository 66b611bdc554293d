"""MI355X rod-dynamics hot path: Python side of librdk.

Mirrors the reference module ``rod_dynamics_py`` (src/python/rod_dynamics_py.cpp);
compute runs in librdk.so (HIP, gfx950) through its C ABI.
"""
from . import _lib
from ._lib import Context, RdkError, lib
from .simulator import Simulator

__all__ = ["Context", "RdkError", "Simulator", "lib", "_lib"]
