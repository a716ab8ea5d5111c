"""Expressions of (t, x, y) for initial and boundary conditions: the
`Expressions` of multigrid/src/input.rs:165-190 evaluated as
multigrid/src/eval/evalexpr.rs does, compiled by the library (sbp_expr_create)
into a program that runs on the device (Eval edges at every stage time,
initial conditions, System::error) or on the host (eval_host).

    e = Expressions.pressure(globals="Mach = 0.4; p0 = 1/(GAMMA * Mach^2);",
                             rho="1", u="...", v="...", p="...")
    sys.set_boundary(block, NORTH, multigrid.Eval(e))
"""
import ctypes

import numpy as np

from . import _lib

CONSERVATION, PRESSURE = _lib.EXPR_CONSERVATION, _lib.EXPR_PRESSURE


class Expressions:
    """A compiled expression set (conservation form: rho, rhou, rhov, e;
    pressure form: rho, u, v, p). GAMMA must be set first (the reference's
    default_context reads it)."""

    def __init__(self, form, globals_, exprs):
        if form not in (CONSERVATION, PRESSURE) or len(exprs) != 4:
            raise ValueError("Expressions(form, globals, [four expressions])")
        self.form = form
        self.globals = globals_ or ""
        self.exprs = list(exprs)
        arr = (ctypes.c_char_p * 4)(*[e.encode() for e in self.exprs])
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        st = _lib.lib().sbp_expr_create(form, self.globals.encode(), arr, ctypes.byref(h), err, len(err))
        if st != _lib.OK:
            raise _lib.SbpError(st, f"sbp_expr_create: {err.value.decode(errors='replace')}")
        self._h = h

    @classmethod
    def conservation(cls, rho, rhou, rhov, e, globals=None):
        return cls(CONSERVATION, globals, [rho, rhou, rhov, e])

    @classmethod
    def pressure(cls, rho, u, v, p, globals=None):
        return cls(PRESSURE, globals, [rho, u, v, p])

    @property
    def handle(self):
        return self._h

    def eval_host(self, t, x, y):
        """(4,) + x.shape conservative variables at time t on the host."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        if x.shape != y.shape:
            raise ValueError("x and y differ in shape")
        out = np.empty((4, x.size))
        dp = ctypes.POINTER(ctypes.c_double)
        _lib.check(_lib.lib().sbp_expr_eval_host(self._h, t, x.size, x.ctypes.data_as(dp), y.ctypes.data_as(dp),
                                                 out.ctypes.data_as(dp)), "sbp_expr_eval_host")
        return out.reshape((4,) + x.shape)

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().sbp_expr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["Expressions", "CONSERVATION", "PRESSURE"]
