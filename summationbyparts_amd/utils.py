"""sbp::utils (sbp/src/utils.rs): Direction and h2linspace."""
import ctypes
from dataclasses import dataclass
from typing import Any

import numpy as np

from . import _lib


@dataclass
class Direction:
    """sbp/src/utils.rs:15-129: one value per block edge."""
    north: Any = None
    south: Any = None
    east: Any = None
    west: Any = None

    def as_tuple(self):
        """(north, south, east, west): the C ABI's sbp_direction order."""
        return (self.north, self.south, self.east, self.west)


def h2linspace(start, end, n):
    """sbp::utils::h2linspace (sbp/src/utils.rs:133-140): start, start + h/2,
    start + 3h/2, ..., end with h = (end - start)/(n - 2)."""
    out = np.empty(n)
    _lib.check(_lib.lib().sbp_h2linspace(start, end, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
               "sbp_h2linspace")
    return out


__all__ = ["Direction", "h2linspace"]
