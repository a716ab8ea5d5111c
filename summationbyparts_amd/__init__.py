"""summationbyparts_amd — MI355X (gfx950) HIP implementation of the
SummationByParts Euler RK4 time-step path.

The compute path is libsbpgpu.so (C ABI in include/sbpgpu.h); this package is
the host-side mirror of the reference's operator and solver interfaces.
"""
from . import _lib
from ._lib import SbpError
from .operators import SBP4, SBP8, OperatorPair, Upwind4, Upwind4h2, Upwind9, Upwind9h2

__all__ = ["SbpError", "SBP4", "SBP8", "OperatorPair", "Upwind4", "Upwind4h2", "Upwind9", "Upwind9h2", "lib"]


def lib():
    """The loaded HIP library (raises ImportError when it is not built)."""
    return _lib.lib()
