"""ctypes wrapper of the oracle (oracle/libsbp_oracle.so) — TEST INFRASTRUCTURE ONLY.

The oracle restates the reference's Euler RK4 path on the CPU (see
sbp_oracle.cpp for the file:line map). Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may import this module; the product package
never does.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libsbp_oracle.so")

UPWIND4, UPWIND4H2, UPWIND9, UPWIND9H2, SBP4, SBP8 = range(6)
OPERATORS = {"upwind4": UPWIND4, "upwind4h2": UPWIND4H2, "upwind9": UPWIND9, "upwind9h2": UPWIND9H2,
             "sbp4": SBP4, "sbp8": SBP8}
DIFF, DISS = 0, 1
XI, ETA = 0, 1
BC_THIS, BC_GRID, BC_VORTEX, BC_INTERPOLATE, BC_MULTIGRID = 0, 1, 2, 3, 4
INTERP4, INTERP8, INTERP9, INTERP9H2 = range(4)
NORTH, SOUTH, EAST, WEST = 0, 1, 2, 3
RHS_AUTO, RHS_TRAD, RHS_UPWIND = 0, 1, 2


class VortexParams(ctypes.Structure):
    _fields_ = [("nvortices", ctypes.c_int32), ("pad_", ctypes.c_int32),
                ("v", (ctypes.c_double * 4) * 5), ("mach", ctypes.c_double)]

    @staticmethod
    def make(vortices, mach):
        p = VortexParams()
        p.nvortices = len(vortices)
        for i, (x0, y0, rstar, eps) in enumerate(vortices):
            p.v[i][0], p.v[i][1], p.v[i][2], p.v[i][3] = x0, y0, rstar, eps
        p.mach = mach
        return p


class OracleError(RuntimeError):
    pass


# Eval boundary data callback: (t, n, x, y, out 4 x n, user)
EVAL_FN = ctypes.CFUNCTYPE(None, ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                           ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)


def _load():
    if not os.path.exists(_LIB_PATH):
        raise OracleError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
    lib = ctypes.CDLL(_LIB_PATH)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    lib.orc_op_info.argtypes = [ctypes.c_int, ip, ip, ip, ip, ip]
    lib.orc_op_h.argtypes = [ctypes.c_int, dp]
    lib.orc_op_block.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp]
    lib.orc_diff.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_int, ctypes.c_int]
    lib.orc_set_gamma.argtypes = [ctypes.c_double]
    lib.orc_set_gamma.restype = None
    lib.orc_vortex.argtypes = [ctypes.POINTER(VortexParams), ctypes.c_double, dp, dp, ctypes.c_long, dp]
    lib.orc_metrics.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_int, ctypes.c_int, dp]
    lib.orc_h2_err.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_int, ctypes.c_int, dp]
    lib.orc_system_create.argtypes = [ctypes.c_int, ip, ip, ip, ip, ctypes.POINTER(dp), ctypes.POINTER(dp), ip]
    lib.orc_system_create.restype = ctypes.c_void_p
    lib.orc_system_destroy.argtypes = [ctypes.c_void_p]
    lib.orc_system_destroy.restype = None
    lib.orc_system_set_bc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(VortexParams)]
    lib.orc_system_set_rhs_kind.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.orc_system_metrics.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
    lib.orc_system_set_state.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
    lib.orc_system_get_state.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
    lib.orc_system_vortex.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(VortexParams)]
    lib.orc_system_rhs.argtypes = [ctypes.c_void_p, ctypes.c_double, dp]
    lib.orc_system_advance.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_long, ctypes.c_int]
    lib.orc_system_time.argtypes = [ctypes.c_void_p]
    lib.orc_system_time.restype = ctypes.c_double
    lib.orc_system_set_time.argtypes = [ctypes.c_void_p, ctypes.c_double]
    lib.orc_system_set_time.restype = None
    lib.orc_system_max_dt.argtypes = [ctypes.c_void_p, dp]
    lib.orc_system_error_vortex.argtypes = [ctypes.c_void_p, ctypes.POINTER(VortexParams), dp]
    lib.orc_system_set_bc_interpolate.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.orc_system_set_bc_multigrid.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ip]
    lib.orc_interpolate.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, dp, ctypes.c_int]
    lib.orc_interp_info.argtypes = [ctypes.c_int, ctypes.c_int, ip]
    lib.orc_interp_coefficients.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
    lib.orc_system_edge.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, dp]
    lib.orc_system_embedded_step.argtypes = [ctypes.c_void_p, dp, ctypes.c_double, dp, ctypes.c_int]
    lib.orc_system_advance_adaptive.argtypes = [ctypes.c_void_p, ctypes.c_double, dp, ctypes.c_double, ctypes.c_int]
    lib.orc_system_set_bc_eval.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, EVAL_FN, ctypes.c_void_p]
    lib.orc_system_set_sat.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.orc_set_inner_threads.argtypes = [ctypes.c_int]
    lib.orc_sat_edge.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, dp, dp]
    lib.orc_system_integrate.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp, dp, ctypes.c_double, ctypes.c_long,
                                         ctypes.c_int]
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
        _lib.orc_set_gamma(1.4)
    return _lib


def _ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(st, what):
    if st != 0:
        raise OracleError(f"{what} failed with status {st}")


def op_info(op):
    vals = [ctypes.c_int() for _ in range(5)]
    _check(lib().orc_op_info(op, *[ctypes.byref(v) for v in vals]), "op_info")
    M, N, D, h2, up = (v.value for v in vals)
    return dict(M=M, N=N, D=D, is_h2=bool(h2), is_upwind=bool(up))


def op_h(op):
    h = np.zeros(op_info(op)["M"])
    _check(lib().orc_op_h(op, _ptr(h)), "op_h")
    return h


def op_block(op, kind):
    info = op_info(op)
    M, N, D = info["M"], info["N"], info["D"]
    start, diag, end = np.zeros((M, N)), np.zeros(D), np.zeros((M, N))
    _check(lib().orc_op_block(op, kind, _ptr(start), _ptr(diag), _ptr(end)), "op_block")
    return start, diag, end


def diff(op, kind, axis, prev):
    prev = np.ascontiguousarray(prev, dtype=np.float64)
    if prev.ndim == 1:
        prev2 = prev[None, :] if axis == XI else prev[:, None]
        return diff(op, kind, axis, np.ascontiguousarray(prev2)).reshape(prev.shape)
    fut = np.zeros_like(prev)
    _check(lib().orc_diff(op, kind, axis, _ptr(prev), _ptr(fut), prev.shape[0], prev.shape[1]), "diff")
    return fut


def vortex(params, t, x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros((4,) + x.shape)
    _check(lib().orc_vortex(ctypes.byref(params), t, _ptr(x), _ptr(y), x.size, _ptr(out)), "vortex")
    return out


def metrics(op_eta, op_xi, x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros((5,) + x.shape)
    _check(lib().orc_metrics(op_eta, op_xi, _ptr(x), _ptr(y), x.shape[0], x.shape[1], _ptr(out)), "metrics")
    return out


def h2_err(op_eta, op_xi, a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = ctypes.c_double()
    _check(lib().orc_h2_err(op_eta, op_xi, _ptr(a), _ptr(b), a.shape[1], a.shape[2], ctypes.byref(out)), "h2_err")
    return out.value


def set_inner_threads(n):
    """Threads inside each block (1, the default, is the reference's one
    thread per block); results do not depend on it."""
    _check(lib().orc_set_inner_threads(int(n)), "set_inner_threads")


def sat_edge(op_eta, op_xi, direction, k, y, metrics, g):
    """SAT_north / south / east / west (euler/src/lib.rs:1025-1144): returns
    k + SAT of one edge; k, y (4, ny, nx), metrics (5, ny, nx), g (4, n)."""
    k = np.array(k, dtype=np.float64, order="C")
    y = np.ascontiguousarray(y, dtype=np.float64)
    metrics = np.ascontiguousarray(metrics, dtype=np.float64)
    g = np.ascontiguousarray(g, dtype=np.float64)
    _check(lib().orc_sat_edge(op_eta, op_xi, direction, y.shape[1], y.shape[2], _ptr(k), _ptr(y), _ptr(metrics), _ptr(g)),
           "sat_edge")
    return k


def interpolate(iop, fine2coarse, data, nout):
    """InterpolationOperator::fine2coarse (fine2coarse=True) or coarse2fine of
    a 1-D array into nout points."""
    data = np.ascontiguousarray(data, dtype=np.float64)
    out = np.zeros(nout)
    _check(lib().orc_interpolate(iop, int(fine2coarse), _ptr(data), data.size, _ptr(out), nout), "interpolate")
    return out


def interp_tables(iop, fine2coarse):
    """(block, diag, (jump0, jump1)) of one direction of an interpolation operator."""
    shape = (ctypes.c_int * 6)()
    _check(lib().orc_interp_info(iop, int(fine2coarse), shape), "interp_info")
    br, bc, dr, dc, j0, j1 = list(shape)
    block, diag = np.zeros((br, bc)), np.zeros((dr, dc))
    _check(lib().orc_interp_coefficients(iop, int(fine2coarse), _ptr(block), _ptr(diag)), "interp_coefficients")
    return block, diag, (j0, j1)


class System:
    """Multi-block system (multigrid SingleThreadedSystem semantics)."""

    def __init__(self, grids, ops):
        """grids: list of (x, y) arrays of shape (ny, nx); ops: list of (op_eta, op_xi)."""
        self.shapes = [np.asarray(x).shape for x, _ in grids]
        nb = len(grids)
        self._keep = [(np.ascontiguousarray(x, dtype=np.float64), np.ascontiguousarray(y, dtype=np.float64))
                      for x, y in grids]
        ny = (ctypes.c_int * nb)(*[s[0] for s in self.shapes])
        nx = (ctypes.c_int * nb)(*[s[1] for s in self.shapes])
        oe = (ctypes.c_int * nb)(*[o[0] for o in ops])
        ox = (ctypes.c_int * nb)(*[o[1] for o in ops])
        xs = (ctypes.POINTER(ctypes.c_double) * nb)(*[_ptr(x) for x, _ in self._keep])
        ys = (ctypes.POINTER(ctypes.c_double) * nb)(*[_ptr(y) for _, y in self._keep])
        st = ctypes.c_int()
        self.h = lib().orc_system_create(nb, ny, nx, oe, ox, xs, ys, ctypes.byref(st))
        if not self.h:
            raise OracleError(f"system_create failed with status {st.value}")
        self.ops = ops

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_system_destroy(self.h)
            self.h = None

    def set_bc(self, block, direction, kind, neighbour=0, vortex_params=None):
        vp = ctypes.byref(vortex_params) if vortex_params is not None else None
        _check(lib().orc_system_set_bc(self.h, block, direction, kind, neighbour, vp), "set_bc")

    def set_bc_interpolate(self, block, direction, neighbour, iop):
        _check(lib().orc_system_set_bc_interpolate(self.h, block, direction, neighbour, iop), "set_bc_interpolate")

    def set_bc_multigrid(self, block, direction, pieces):
        """pieces: list of (block, start, end)."""
        flat = (ctypes.c_int * (3 * len(pieces)))(*[v for p in pieces for v in p])
        _check(lib().orc_system_set_bc_multigrid(self.h, block, direction, len(pieces), flat), "set_bc_multigrid")

    def edge(self, block, direction, t=0.0):
        """Boundary data (4, n) of one edge for the current state."""
        ny, nx = self.shapes[block]
        n = nx if direction in (NORTH, SOUTH) else ny
        out = np.zeros((4, n))
        _check(lib().orc_system_edge(self.h, block, direction, t, _ptr(out)), "edge")
        return out

    def set_rhs_kind(self, kind):
        _check(lib().orc_system_set_rhs_kind(self.h, kind), "set_rhs_kind")

    def metrics(self, block):
        out = np.zeros((5,) + self.shapes[block])
        _check(lib().orc_system_metrics(self.h, block, _ptr(out)), "metrics")
        return out

    def set_state(self, block, state):
        state = np.ascontiguousarray(state, dtype=np.float64)
        assert state.shape == (4,) + self.shapes[block]
        _check(lib().orc_system_set_state(self.h, block, _ptr(state)), "set_state")

    def get_state(self, block):
        out = np.zeros((4,) + self.shapes[block])
        _check(lib().orc_system_get_state(self.h, block, _ptr(out)), "get_state")
        return out

    def vortex(self, t, params):
        _check(lib().orc_system_vortex(self.h, t, ctypes.byref(params)), "vortex")

    def rhs(self, t):
        sizes = [4 * s[0] * s[1] for s in self.shapes]
        out = np.zeros(sum(sizes))
        _check(lib().orc_system_rhs(self.h, t, _ptr(out)), "rhs")
        res, off = [], 0
        for s, n in zip(self.shapes, sizes):
            res.append(out[off:off + n].reshape((4,) + s))
            off += n
        return res

    def set_bc_eval(self, block, direction, evaluator):
        """BoundaryCharacteristic::Eval (euler/src/lib.rs:741): `evaluator`
        maps (t, x, y) arrays of edge points to (rho, rhou, rhov, e) arrays
        (e.g. oracle.evalexpr.Evaluator.field)."""
        def cb(t, n, x, y, out, _user):
            X = np.ctypeslib.as_array(x, shape=(n,)).copy()
            Y = np.ctypeslib.as_array(y, shape=(n,)).copy()
            q = np.asarray(evaluator(t, X, Y), dtype=np.float64).reshape(4, n)
            np.ctypeslib.as_array(out, shape=(4 * n,))[:] = q.ravel()
        fn = EVAL_FN(cb)
        self._callbacks = getattr(self, "_callbacks", []) + [fn]  # kept alive with the system
        _check(lib().orc_system_set_bc_eval(self.h, block, direction, fn, None), "set_bc_eval")

    def set_sat(self, on):
        """False: the right-hand side is RHS_no_SAT (euler/src/lib.rs:518-568)."""
        _check(lib().orc_system_set_sat(self.h, int(bool(on))), "set_sat")

    def integrate(self, A, B, C, dt, nsteps, nthreads=1):
        """nsteps of integrate::<tableau> (integrate/src/lib.rs:176-222):
        A (s, s) with stage i using A[i, :i], B (s), C (s-1 used)."""
        A = np.ascontiguousarray(A, dtype=np.float64)
        B = np.ascontiguousarray(B, dtype=np.float64)
        C = np.ascontiguousarray(np.concatenate([np.asarray(C, dtype=np.float64), [0.0]]))
        _check(lib().orc_system_integrate(self.h, len(B), _ptr(A), _ptr(B), _ptr(C), dt, nsteps, nthreads), "integrate")

    def advance(self, dt, nsteps, nthreads=1):
        _check(lib().orc_system_advance(self.h, dt, nsteps, nthreads), "advance")

    @property
    def time(self):
        return lib().orc_system_time(self.h)

    @time.setter
    def time(self, t):
        lib().orc_system_set_time(self.h, t)

    def embedded_step(self, time, dt, nthreads=1):
        """One BogackiShampine step (integrate_embedded_rk) at loop time `time`;
        the state becomes fut; returns (time + dt, fut2 per block)."""
        t = ctypes.c_double(time)
        sizes = [4 * s[0] * s[1] for s in self.shapes]
        out = np.zeros(sum(sizes))
        _check(lib().orc_system_embedded_step(self.h, ctypes.byref(t), dt, _ptr(out), nthreads), "embedded_step")
        res, off = [], 0
        for s, n in zip(self.shapes, sizes):
            res.append(out[off:off + n].reshape((4,) + s))
            off += n
        return t.value, res

    def advance_adaptive(self, dt, guess_dt, maxerr, nthreads=1):
        """euler::System::advance_adaptive; returns the updated guess."""
        g = ctypes.c_double(guess_dt)
        _check(lib().orc_system_advance_adaptive(self.h, dt, ctypes.byref(g), maxerr, nthreads), "advance_adaptive")
        return g.value

    def max_dt(self):
        out = ctypes.c_double()
        _check(lib().orc_system_max_dt(self.h, ctypes.byref(out)), "max_dt")
        return out.value

    def error_vortex(self, params):
        out = ctypes.c_double()
        _check(lib().orc_system_error_vortex(self.h, ctypes.byref(params), ctypes.byref(out)), "error")
        return out.value
