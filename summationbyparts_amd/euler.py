"""Compressible Euler: the `euler` crate's surface (euler/src/lib.rs,
euler/src/vortex.rs) over the fused HIP RK4 path.

    set_gamma(1.4)                                   # euler::GAMMA.set
    sys = System(x, y, Upwind9())                    # euler::System::new  (lib.rs:29-48)
    sys.vortex(0.0, VortexParameters([Vortice(-1, 0, 0.5, 1)], mach=0.5))
    sys.advance_upwind(dt)                           # lib.rs:124-148 (periodic This BCs)
    q = sys.field()                                  # (4, ny, nx): rho, rhou, rhov, e
"""
import ctypes
from dataclasses import dataclass, field as dc_field
from typing import List

import numpy as np

from . import _lib
from .operators import OperatorPair


def set_gamma(gamma=1.4):
    """euler::GAMMA.set (a OnceCell: setting another value later is an error)."""
    _lib.check(_lib.lib().sbp_set_gamma(gamma), "sbp_set_gamma")


def gamma():
    g = ctypes.c_double()
    _lib.check(_lib.lib().sbp_get_gamma(ctypes.byref(g)), "sbp_get_gamma")
    return g.value


@dataclass
class Vortice:
    """euler/src/vortex.rs:5-10"""
    x0: float
    y0: float
    rstar: float
    eps: float


@dataclass
class VortexParameters:
    """euler/src/vortex.rs:17-21 (at most 5 vortices)."""
    vortices: List[Vortice] = dc_field(default_factory=list)
    mach: float = 0.5

    def c(self):
        return _lib.vortex_params_c([(v.x0, v.y0, v.rstar, v.eps) for v in self.vortices], self.mach)


def _ids(op):
    return op.ids


class BlockSystem:
    """Blocks stepped together by the fused RK4 kernels (shared base of
    euler.System and multigrid.System)."""

    def __init__(self, grids, ops):
        self._ctx = _lib.context()
        nb = len(grids)
        self.shapes = []
        xs, ys = [], []
        for x, y in grids:
            x = np.ascontiguousarray(x, dtype=np.float64)
            y = np.ascontiguousarray(y, dtype=np.float64)
            if x.shape != y.shape or x.ndim != 2:
                raise ValueError("x and y must be 2-D arrays of one shape")
            self.shapes.append(x.shape)
            xs.append(x)
            ys.append(y)
        self._grids = list(zip(xs, ys))
        self.ops = list(ops)
        ids = [_ids(o) for o in self.ops]
        dp = ctypes.POINTER(ctypes.c_double)
        ny = (ctypes.c_int * nb)(*[s[0] for s in self.shapes])
        nx = (ctypes.c_int * nb)(*[s[1] for s in self.shapes])
        oe = (ctypes.c_int * nb)(*[i[0] for i in ids])
        ox = (ctypes.c_int * nb)(*[i[1] for i in ids])
        xp = (dp * nb)(*[x.ctypes.data_as(dp) for x in xs])
        yp = (dp * nb)(*[y.ctypes.data_as(dp) for y in ys])
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().sbp_system_create(self._ctx, nb, ny, nx, oe, ox, xp, yp, ctypes.byref(h)),
                   "sbp_system_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().sbp_system_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- configuration
    def set_bc(self, block, direction, kind, neighbour=0, vortex=None):
        vp = ctypes.byref(vortex.c()) if vortex is not None else None
        _lib.check(_lib.lib().sbp_system_set_bc(self._h, block, direction, kind, neighbour, vp), "sbp_system_set_bc")

    def set_bc_interpolate(self, block, direction, neighbour, interp_op):
        """BoundaryCharacteristic::Interpolate(neighbour, operator) (euler/src/lib.rs:743)."""
        _lib.check(_lib.lib().sbp_system_set_bc_interpolate(self._h, block, direction, neighbour, interp_op),
                   "sbp_system_set_bc_interpolate")

    def set_bc_multigrid(self, block, direction, pieces):
        """BoundaryCharacteristic::MultiGrid(pieces) (euler/src/lib.rs:744): (block, start, end) triples."""
        flat = (ctypes.c_int * (3 * len(pieces)))(*[int(v) for p in pieces for v in p])
        _lib.check(_lib.lib().sbp_system_set_bc_multigrid(self._h, block, direction, len(pieces), flat),
                   "sbp_system_set_bc_multigrid")

    def set_rhs_kind(self, kind):
        _lib.check(_lib.lib().sbp_system_set_rhs_kind(self._h, kind), "sbp_system_set_rhs_kind")

    def set_precision(self, precision):
        _lib.check(_lib.lib().sbp_system_set_precision(self._h, precision), "sbp_system_set_precision")

    def set_tile_shape(self, index):
        _lib.check(_lib.lib().sbp_system_set_tile_shape(self._h, index), "sbp_system_set_tile_shape")

    def tile_shape(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.lib().sbp_system_tile_shape(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
                   "sbp_system_tile_shape")
        return a.value, b.value, c.value

    def tile_index(self):
        i = ctypes.c_int()
        _lib.check(_lib.lib().sbp_system_tile_index(self._h, ctypes.byref(i)), "sbp_system_tile_index")
        return i.value

    def set_schedule(self, schedule):
        """Stage schedule: 0 split over two streams, 1 one dispatch over all
        tiles, 2 the edge SATs by a separate kernel and then one dispatch,
        -1 autotuned (results do not depend on it)."""
        _lib.check(_lib.lib().sbp_system_set_schedule(self._h, schedule), "sbp_system_set_schedule")

    def schedule(self):
        i = ctypes.c_int()
        _lib.check(_lib.lib().sbp_system_schedule(self._h, ctypes.byref(i)), "sbp_system_schedule")
        return i.value

    def set_tile_order(self, order):
        """Tile order of the dispatches: 0 XCD bands, 1 row-major, -1 autotuned
        (results do not depend on it)."""
        _lib.check(_lib.lib().sbp_system_set_tile_order(self._h, order), "sbp_system_set_tile_order")

    def tile_order(self):
        """Order of the tiles in a dispatch: 0 XCD bands, 1 row-major (tuned with the schedule)."""
        i = ctypes.c_int()
        _lib.check(_lib.lib().sbp_system_tile_order(self._h, ctypes.byref(i)), "sbp_system_tile_order")
        return i.value

    # ---- state
    def set_state(self, block, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        if q.shape != (4,) + self.shapes[block]:
            raise ValueError(f"state must be (4, ny, nx) = {(4,) + self.shapes[block]}")
        _lib.check(_lib.lib().sbp_system_set_state(self._h, block, q.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                   "sbp_system_set_state")

    def get_state(self, block):
        out = np.zeros((4,) + self.shapes[block])
        _lib.check(_lib.lib().sbp_system_get_state(self._h, block, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                   "sbp_system_get_state")
        return out

    def metrics(self, block):
        out = np.zeros((5,) + self.shapes[block])
        _lib.check(_lib.lib().sbp_system_get_metrics(self._h, block, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                   "sbp_system_get_metrics")
        return out

    def vortex(self, t, params, block=None):
        """Field::vortex on every block, or on one block."""
        p = params.c()
        if block is None:
            _lib.check(_lib.lib().sbp_system_vortex(self._h, t, ctypes.byref(p)), "sbp_system_vortex")
        else:
            _lib.check(_lib.lib().sbp_system_vortex_block(self._h, block, t, ctypes.byref(p)), "sbp_system_vortex_block")

    def set_dt_scale(self, block, scale):
        """Block `block` steps with dt·scale; its time is t·scale."""
        _lib.check(_lib.lib().sbp_system_set_dt_scale(self._h, block, scale), "sbp_system_set_dt_scale")

    # ---- stepping
    def rhs(self, t):
        sizes = [4 * s[0] * s[1] for s in self.shapes]
        out = np.zeros(sum(sizes))
        _lib.check(_lib.lib().sbp_system_rhs(self._h, t, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                   "sbp_system_rhs")
        res, off = [], 0
        for s, n in zip(self.shapes, sizes):
            res.append(out[off:off + n].reshape((4,) + s))
            off += n
        return res

    def advance_steps(self, dt, nsteps):
        _lib.check(_lib.lib().sbp_system_advance(self._h, dt, nsteps), "sbp_system_advance")

    @property
    def time(self):
        t = ctypes.c_double()
        _lib.check(_lib.lib().sbp_system_get_time(self._h, ctypes.byref(t)), "sbp_system_get_time")
        return t.value

    @time.setter
    def time(self, t):
        _lib.check(_lib.lib().sbp_system_set_time(self._h, t), "sbp_system_set_time")

    def max_dt(self):
        out = ctypes.c_double()
        _lib.check(_lib.lib().sbp_system_max_dt(self._h, ctypes.byref(out)), "sbp_system_max_dt")
        return out.value

    def error_vortex(self, params, block=None):
        """Σ over blocks (or one block) of Field::h2_err against the vortex."""
        out = ctypes.c_double()
        p = params.c()
        if block is None:
            _lib.check(_lib.lib().sbp_system_error_vortex(self._h, ctypes.byref(p), ctypes.byref(out)),
                       "sbp_system_error_vortex")
        else:
            _lib.check(_lib.lib().sbp_system_error_vortex_block(self._h, block, ctypes.byref(p), ctypes.byref(out)),
                       "sbp_system_error_vortex_block")
        return out.value

    def energy(self):
        out = ctypes.c_double()
        _lib.check(_lib.lib().sbp_system_energy(self._h, ctypes.byref(out)), "sbp_system_energy")
        return out.value

    def advance_adaptive_steps(self, dt, guess_dt, maxerr):
        """sbp_system_advance_adaptive with the system's boundaries and RHS
        kind; returns the updated guess."""
        g = ctypes.c_double(guess_dt)
        _lib.check(_lib.lib().sbp_system_advance_adaptive(self._h, dt, ctypes.byref(g), maxerr),
                   "sbp_system_advance_adaptive")
        return g.value

    def set_bc_eval(self, block, direction, expressions):
        """BoundaryCharacteristic::Eval (lib.rs:741) on one edge."""
        _lib.check(_lib.lib().sbp_system_set_bc_eval(self._h, block, direction, expressions.handle),
                   "sbp_system_set_bc_eval")

    def eval_block(self, block, t, expressions):
        """The state of one block from expressions at time t (on the device)."""
        _lib.check(_lib.lib().sbp_system_eval_block(self._h, block, t, expressions.handle), "sbp_system_eval_block")

    def error_eval(self, expressions):
        """System::error against expressions at the system time."""
        out = ctypes.c_double()
        _lib.check(_lib.lib().sbp_system_error_eval(self._h, expressions.handle, ctypes.byref(out)),
                   "sbp_system_error_eval")
        return out.value

    def integrate(self, tableau, dt, nsteps=1):
        """nsteps of integrate::<tableau> (utils/integrate/src/lib.rs:176-222)
        for any explicit tableau (summationbyparts_amd.integrate.Rk4, Rk4_38,
        EulerMethod, MidpointMethod, Rk6, Fehlberg, BogackiShampine, or an
        object with A, B, C in the reference's form)."""
        from .integrate import tableau_arrays
        ns, A, B, C = tableau_arrays(tableau)
        dp = ctypes.POINTER(ctypes.c_double)
        _lib.check(_lib.lib().sbp_system_integrate(self._h, ns, A.ctypes.data_as(dp), B.ctypes.data_as(dp),
                                                   C.ctypes.data_as(dp), dt, nsteps), "sbp_system_integrate")

    def advance_timed(self, dt, nsteps):
        """advance_steps(dt, nsteps); returns the ms between HIP events
        recorded on the library's stream around the steps."""
        ms = ctypes.c_double()
        _lib.check(_lib.lib().sbp_system_advance_timed(self._h, dt, nsteps, ctypes.byref(ms)),
                   "sbp_system_advance_timed")
        return ms.value

    def profile(self, dt, nsteps):
        """(ms for nsteps graph-replayed steps, mean ms of one stage-kernel launch)."""
        tot, avg = ctypes.c_double(), ctypes.c_double()
        _lib.check(_lib.lib().sbp_system_profile(self._h, dt, nsteps, ctypes.byref(tot), ctypes.byref(avg)),
                   "sbp_system_profile")
        return tot.value, avg.value


class System(BlockSystem):
    """euler::System (euler/src/lib.rs:20-186): one block, periodic (`This`)
    boundaries on all four sides."""

    def __init__(self, x, y, op):
        super().__init__([(x, y)], [op])
        self.op = op
        self._x, self._y = self._grids[0]

    def advance(self, dt):
        """RK4 step with RHS_trad (lib.rs:50-74)."""
        self.set_rhs_kind(_lib.RHS_TRAD)
        self.advance_steps(dt, 1)

    def advance_upwind(self, dt):
        """RK4 step with RHS_upwind (lib.rs:124-148)."""
        if self.op.upwind() is None:
            raise _lib.SbpError(_lib.ERR_NOT_UPWIND, "advance_upwind")
        self.set_rhs_kind(_lib.RHS_UPWIND)
        self.advance_steps(dt, 1)

    def advance_adaptive(self, dt, guess_dt, maxerr):
        """euler::System::advance_adaptive (lib.rs:149-185): embedded
        Bogacki-Shampine steps with RHS_upwind until the reference's loop time
        reaches dt; returns the updated guess (the reference's `&mut guess_dt`)."""
        if self.op.upwind() is None:
            raise _lib.SbpError(_lib.ERR_NOT_UPWIND, "advance_adaptive")
        self.set_rhs_kind(_lib.RHS_UPWIND)
        return self.advance_adaptive_steps(dt, guess_dt, maxerr)

    def advance_n(self, dt, nsteps, upwind=True):
        """nsteps RK4 steps in one call (graph-replayed)."""
        self.set_rhs_kind(_lib.RHS_UPWIND if upwind else _lib.RHS_TRAD)
        self.advance_steps(dt, nsteps)

    def field(self):
        return self.get_state(0)

    def x(self):
        return self._x

    def y(self):
        return self._y

    def nx(self):
        return self.shapes[0][1]

    def ny(self):
        return self.shapes[0][0]


def rhs(kind, op, k, y, metrics, boundaries):
    """euler::RHS_trad / RHS_upwind (lib.rs:483-515, :571-616) on device
    tensors: k, y (4, ny, nx); metrics (5, ny, nx); boundaries = (north,
    south, east, west) boundary data, each (4, n) contiguous."""
    import torch

    for t in (k, y, metrics) + tuple(boundaries):
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
            raise ValueError("rhs needs contiguous float64 CUDA tensors")
    ny, nx = y.shape[1], y.shape[2]
    oe, ox = _ids(op)
    bp = (ctypes.c_void_p * 4)(*[b.data_ptr() for b in boundaries])
    ctx = _lib.context()
    _lib.wait_for_torch()
    _lib.check(_lib.lib().sbp_rhs(ctx, kind, oe, ox, ny, nx, ctypes.c_void_p(k.data_ptr()),
                                  ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(metrics.data_ptr()), bp), "sbp_rhs")


def RHS_upwind(op, k, y, metrics, boundaries):
    rhs(_lib.RHS_UPWIND, op, k, y, metrics, boundaries)


def RHS_trad(op, k, y, metrics, boundaries):
    rhs(_lib.RHS_TRAD, op, k, y, metrics, boundaries)


def _check_fields(*ts):
    import torch

    for t in ts:
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
            raise ValueError("need contiguous float64 CUDA tensors")


def RHS_no_SAT(op, k, y, metrics):
    """euler::RHS_no_SAT (lib.rs:518-568): the right-hand side without
    boundary penalties; upwind dissipation iff op.upwind() is not None."""
    _check_fields(k, y, metrics)
    oe, ox = _ids(op)
    _lib.wait_for_torch()
    _lib.check(_lib.lib().sbp_rhs_no_sat(_lib.context(), oe, ox, y.shape[1], y.shape[2], ctypes.c_void_p(k.data_ptr()),
                                         ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(metrics.data_ptr())),
               "sbp_rhs_no_sat")


def SAT_characteristics(op, k, y, metrics, boundaries):
    """euler::SAT_characteristics (lib.rs:1002-1013): k += the SATs of the
    north, south, east and west edges; boundaries = (north, south, east,
    west) boundary data, each (4, n) contiguous."""
    _check_fields(k, y, metrics, *boundaries)
    oe, ox = _ids(op)
    bp = (ctypes.c_void_p * 4)(*[b.data_ptr() for b in boundaries])
    _lib.wait_for_torch()
    _lib.check(_lib.lib().sbp_sat_characteristics(_lib.context(), oe, ox, y.shape[1], y.shape[2],
                                                  ctypes.c_void_p(k.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                                  ctypes.c_void_p(metrics.data_ptr()), bp), "sbp_sat_characteristics")


def _sat_edge(direction):
    def sat(op, k, y, metrics, boundary):
        _check_fields(k, y, metrics, boundary)
        oe, ox = _ids(op)
        _lib.wait_for_torch()
        _lib.check(_lib.lib().sbp_sat_edge(_lib.context(), direction, oe, ox, y.shape[1], y.shape[2],
                                           ctypes.c_void_p(k.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                           ctypes.c_void_p(metrics.data_ptr()), ctypes.c_void_p(boundary.data_ptr())),
                   "sbp_sat_edge")
    return sat


SAT_north = _sat_edge(_lib.NORTH)
SAT_south = _sat_edge(_lib.SOUTH)
SAT_east = _sat_edge(_lib.EAST)
SAT_west = _sat_edge(_lib.WEST)
# euler::SAT_FUNCTIONS (lib.rs:1015-1023), keyed by direction
SAT_FUNCTIONS = {"north": SAT_north, "south": SAT_south, "east": SAT_east, "west": SAT_west}


def h2_err(a, b, op):
    """Field::h2_err (lib.rs:377-432) of two (4, ny, nx) device fields."""
    _check_fields(a, b)
    if tuple(a.shape) != tuple(b.shape):
        raise ValueError("h2_err: fields differ in shape")
    oe, ox = _ids(op)
    out = ctypes.c_double()
    _lib.wait_for_torch()
    _lib.check(_lib.lib().sbp_h2_err(_lib.context(), oe, ox, a.shape[1], a.shape[2], ctypes.c_void_p(a.data_ptr()),
                                     ctypes.c_void_p(b.data_ptr()), ctypes.byref(out)), "sbp_h2_err")
    return out.value


def grid_metrics(op, x, y):
    """Grid::metrics (sbp/src/grid.rs:64-66) on device tensors: (5, ny, nx)."""
    import torch

    ny, nx = x.shape
    out = torch.empty((5, ny, nx), dtype=torch.float64, device=x.device)
    oe, ox = _ids(op)
    _lib.wait_for_torch()
    _lib.check(_lib.lib().sbp_metrics(_lib.context(), oe, ox, ny, nx, ctypes.c_void_p(x.data_ptr()),
                                      ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(out.data_ptr())), "sbp_metrics")
    return out


__all__ = ["set_gamma", "gamma", "Vortice", "VortexParameters", "System", "BlockSystem", "RHS_upwind", "RHS_trad",
           "RHS_no_SAT", "SAT_characteristics", "SAT_north", "SAT_south", "SAT_east", "SAT_west", "SAT_FUNCTIONS",
           "h2_err", "grid_metrics", "OperatorPair"]
