"""Butcher tableaus of utils/integrate/src/lib.rs (ButcherTableau :17-24),
stepped on the GPU by BlockSystem.integrate (sbp_system_integrate). A, B, C
keep the reference's form: A[i-1] holds stage i's weights of k_0..k_{i-1}
(row i of a lower-triangular matrix), stage i > 0 runs at t + dt·C[i-1].
RK4 itself is fused (BlockSystem.advance_steps); these tableaus take the
unfused path of one right-hand-side launch per stage.
"""
import math

import numpy as np


class Rk4:
    """integrate/src/lib.rs:30-35"""
    A = ((0.5,), (0.0, 0.5), (0.0, 0.0, 1.0))
    B = (1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0)
    C = (0.5, 0.5, 1.0)


class Rk4_38:
    """integrate/src/lib.rs:37-42"""
    A = ((1.0 / 3.0,), (-1.0 / 3.0, 1.0), (1.0, -1.0, 1.0))
    B = (1.0 / 8.0, 3.0 / 8.0, 3.0 / 8.0, 1.0 / 8.0)
    C = (1.0 / 3.0, 2.0 / 3.0, 1.0)


class EulerMethod:
    """integrate/src/lib.rs:44-49"""
    A = ((),)
    B = (1.0,)
    C = ()


class MidpointMethod:
    """integrate/src/lib.rs:51-56"""
    A = ((0.5,),)
    B = (0.0, 1.0)
    C = (0.5,)


_S5 = math.sqrt(5.0)


class Rk6:
    """integrate/src/lib.rs:58-110 (SQRT_5 there is the same double as math.sqrt(5))"""
    A = ((4.0 / 7.0,),
         (115.0 / 112.0, -5.0 / 16.0),
         (589.0 / 630.0, 5.0 / 18.0, -16.0 / 45.0),
         (229.0 / 1200.0 - 29.0 / 6000.0 * _S5, 119.0 / 240.0 - 187.0 / 1200.0 * _S5, -14.0 / 75.0 + 34.0 / 375.0 * _S5,
          -3.0 / 100.0 * _S5),
         (71.0 / 2400.0 - 587.0 / 12000.0 * _S5, 187.0 / 480.0 - 391.0 / 2400.0 * _S5, -38.0 / 75.0 + 26.0 / 375.0 * _S5,
          27.0 / 80.0 - 3.0 / 400.0 * _S5, (1.0 + _S5) / 4.0),
         (-49.0 / 480.0 + 43.0 / 160.0 * _S5, -425.0 / 96.0 + 51.0 / 32.0 * _S5, 52.0 / 15.0 - 4.0 / 5.0 * _S5,
          -27.0 / 16.0 + 3.0 / 16.0 * _S5, 5.0 / 4.0 - 3.0 / 4.0 * _S5, 5.0 / 2.0 - 1.0 / 2.0 * _S5))
    B = (1.0 / 12.0, 0.0, 0.0, 0.0, 5.0 / 12.0, 5.0 / 12.0, 1.0 / 12.0)
    C = (4.0 / 7.0, 5.0 / 7.0, 6.0 / 7.0, (5.0 - _S5) / 10.0, (5.0 + _S5) / 10.0, 1.0)


class Fehlberg:
    """integrate/src/lib.rs:112-137"""
    A = ((1.0 / 4.0,),
         (3.0 / 32.0, 9.0 / 32.0),
         (1932.0 / 2197.0, -7200.0 / 2197.0, 7296.0 / 2197.0),
         (439.0 / 216.0, -8.0, 3680.0 / 513.0, -845.0 / 4104.0),
         (-8.0 / 27.0, 2.0, -3544.0 / 2565.0, 1859.0 / 4104.0, -11.0 / 40.0))
    B = (16.0 / 135.0, 0.0, 6656.0 / 12825.0, 28561.0 / 56430.0, -9.0 / 50.0, 2.0 / 55.0)
    C = (0.0, 1.0 / 4.0, 3.0 / 8.0, 12.0 / 13.0, 1.0, 1.0 / 2.0)
    BSTAR = (25.0 / 216.0, 0.0, 1408.0 / 2565.0, 2197.0 / 4104.0, -1.0 / 5.0, 0.0)


class BogackiShampine:
    """integrate/src/lib.rs:139-151 (C starts with its 0, so integrate's
    stage i runs at t + dt·C[i-1]: t, t, t + dt/2, t + 3dt/4)"""
    A = ((1.0 / 2.0,), (0.0, 3.0 / 4.0), (2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0))
    B = (2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0, 0.0)
    C = (0.0, 1.0 / 2.0, 3.0 / 4.0, 1.0)
    BSTAR = (7.0 / 24.0, 1.0 / 4.0, 1.0 / 3.0, 1.0 / 8.0)


def tableau_arrays(t):
    """(s, A (s x s row-major), B (s), C (max(s - 1, 1))) as the C ABI takes them."""
    B = np.array(t.B, dtype=np.float64)
    s = len(B)
    A = np.zeros((s, s))
    for i in range(1, s):
        row = t.A[i - 1]
        A[i, :len(row)] = row
    C = np.zeros(max(s - 1, 1))
    C[:s - 1] = np.array(t.C, dtype=np.float64)[:s - 1]
    return s, np.ascontiguousarray(A), B, C


__all__ = ["Rk4", "Rk4_38", "EulerMethod", "MidpointMethod", "Rk6", "Fehlberg", "BogackiShampine", "tableau_arrays"]
