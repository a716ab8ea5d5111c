"""The multigrid front end: a JSON5 configuration (multigrid/src/input.rs) read
into a runtime configuration (multigrid/src/parsing.rs) and a system on the
GPU, driven as multigrid/src/main.rs drives it.

    cfg = Configuration.loads(open("quad.json").read())   # json5::from_str (main.rs:57)
    rt = cfg.into_runtime()                                 # parsing.rs:46-233
    sys = rt.system()                                       # BaseSystem::new + create (system.rs:28-158)
    run(rt, sys)                                            # main.rs:96-147: dt = max_dt, round(T/dt) steps

Configuration errors the reference panics on (a missing operator, an
unknown neighbour name, vortex edges without vortex boundary data, ...)
raise ConfigError here.
"""
import math
from dataclasses import dataclass, field as dc_field
from typing import List, Optional

import numpy as np

from . import _lib, json5, multigrid, operators
from .euler import VortexParameters, Vortice, set_gamma
from .expressions import Expressions
from .utils import Direction, h2linspace

__all__ = ["Configuration", "RuntimeConfiguration", "ConfigError", "run", "load"]


class ConfigError(ValueError):
    pass


# input.rs:6-15 (serde rename_all = "lowercase")
OPERATORS = {"upwind4": operators.Upwind4, "upwind9": operators.Upwind9, "upwind4h2": operators.Upwind4h2,
             "upwind9h2": operators.Upwind9h2, "sbp4": operators.SBP4, "sbp8": operators.SBP8}
# input.rs:109-119
INTERPOLATION = {"4": operators.Interpolation4, "8": operators.Interpolation8, "9": operators.Interpolation9,
                 "9h2": operators.Interpolation9h2}
DIRECTIONS = ("north", "south", "east", "west")


def _one_key(obj, what):
    """An externally tagged enum value ({"variant": payload})."""
    if not isinstance(obj, dict) or len(obj) != 1:
        raise ConfigError(f"{what}: expected an object with one key, got {obj!r}")
    return next(iter(obj.items()))


def _linspace(start, end, steps):
    """ndarray's Array::linspace (start + step·i with step = (end - start)/(n - 1))."""
    if steps < 1:
        raise ConfigError("linspace needs at least one step")
    step = (end - start) / (steps - 1) if steps > 1 else 0.0
    return np.array([start + step * i for i in range(steps)], dtype=np.float64)


def _array_form(v, what):
    """input.rs:31-60: {"linspace": {...}} or {"array": ndarray}; ndarray's
    serde form is {"v": 1, "dim": [...], "data": [...]}."""
    kind, payload = _one_key(v, what)
    if kind == "linspace":
        try:
            start, end, steps = float(payload["start"]), float(payload["end"]), payload["steps"]
        except (KeyError, TypeError) as e:
            raise ConfigError(f"{what}: linspace needs start, end and steps") from e
        if not isinstance(steps, int) or steps < 0:
            raise ConfigError(f"{what}: steps is an unsigned integer")
        return h2linspace(start, end, steps) if payload.get("h2", False) else _linspace(start, end, steps)
    if kind == "array":
        try:
            dim, data = payload["dim"], payload["data"]
        except (KeyError, TypeError) as e:
            raise ConfigError(f"{what}: array needs dim and data") from e
        a = np.array(data, dtype=np.float64)
        if len(dim) not in (1, 2) or a.size != int(np.prod(dim)):
            raise ConfigError(f"{what}: bad array")
        return a.reshape(dim)
    raise ConfigError(f"{what}: unknown grid form {kind!r}")


def _vortex_params(v):
    """euler::VortexParameters (vortex.rs:16-21, serde)."""
    try:
        vs = [Vortice(float(w["x0"]), float(w["y0"]), float(w["rstar"]), float(w["eps"])) for w in v["vortices"]]
        return VortexParameters(vs, float(v["mach"]))
    except (KeyError, TypeError) as e:
        raise ConfigError(f"bad vortex parameters {v!r}") from e


def _expressions(v):
    """input.rs:165-190: untagged, conservation form (rho, rhou, rhov, e)
    first, then the pressure form (rho, u, v, p)."""
    if all(k in v for k in ("rho", "rhou", "rhov", "e")):
        return Expressions.conservation(v["rho"], v["rhou"], v["rhov"], v["e"], globals=v.get("globals"))
    if all(k in v for k in ("rho", "u", "v", "p")):
        return Expressions.pressure(v["rho"], v["u"], v["v"], v["p"], globals=v.get("globals"))
    raise ConfigError(f"expressions need rho, rhou, rhov, e or rho, u, v, p: {sorted(v)}")


@dataclass
class GridConfig:
    """input.rs:151-157 (every field optional; `default` fills the gaps)."""
    operators: Optional[dict] = None
    x: Optional[dict] = None
    y: Optional[dict] = None
    boundary_conditions: Optional[dict] = None


@dataclass
class Configuration:
    """input.rs:221-232."""
    grids: dict
    integration_time: float
    initial_conditions: dict
    boundary_conditions: object = "not_needed"
    gamma: float = 1.4

    @classmethod
    def from_value(cls, v):
        if not isinstance(v, dict):
            raise ConfigError("the configuration is an object")
        for k in ("grids", "integration_time", "initial_conditions"):
            if k not in v:
                raise ConfigError(f"missing field {k!r}")
        grids = {}
        for name, g in v["grids"].items():
            if not isinstance(g, dict):
                raise ConfigError(f"grid {name!r} is not an object")
            unknown = set(g) - {"operators", "x", "y", "boundary_conditions"}
            if unknown:
                raise ConfigError(f"grid {name!r}: unknown fields {sorted(unknown)}")
            grids[name] = GridConfig(g.get("operators"), g.get("x"), g.get("y"), g.get("boundary_conditions"))
        return cls(grids, float(v["integration_time"]), v["initial_conditions"],
                   v.get("boundary_conditions", "not_needed"), float(v.get("gamma", 1.4)))

    @classmethod
    def loads(cls, text):
        return cls.from_value(json5.loads(text))

    def into_runtime(self):
        return _into_runtime(self)


@dataclass
class RuntimeConfiguration:
    """parsing.rs:37-44: grid names, coordinates, boundary conditions per
    edge, operator pairs (eta, xi), integration time, initial conditions
    (a VortexParameters or an Expressions)."""
    names: List[str]
    grids: list
    boundary_conditions: list
    operators: list
    integration_time: float
    initial_conditions: object
    gamma: float = 1.4
    expressions: list = dc_field(default_factory=list)  # keeps the compiled programs alive

    def system(self):
        """BaseSystem::create (system.rs:56-158): the blocks, their boundary
        conditions and the initial conditions at t = 0 on the GPU."""
        sys_ = multigrid.System(self.grids, self.operators)
        for i, bc in enumerate(self.boundary_conditions):
            for d, b in zip((multigrid.NORTH, multigrid.SOUTH, multigrid.EAST, multigrid.WEST), bc.as_tuple()):
                sys_.set_boundary(i, d, b)
        if isinstance(self.initial_conditions, VortexParameters):
            sys_.vortex(0.0, self.initial_conditions)
        else:
            for i in range(len(self.grids)):
                sys_.eval_block(i, 0.0, self.initial_conditions)
        return sys_

    def error(self, sys_):
        """System::error (system.rs:618-635) against the initial conditions at
        the system time."""
        if isinstance(self.initial_conditions, VortexParameters):
            return sys_.error_vortex(self.initial_conditions)
        return sys_.error_eval(self.initial_conditions)


def _broadcast(x, y, name):
    """parsing.rs:88-128: 1-D coordinates broadcast to (ny, nx); a 1-D x with a
    2-D y is broadcast to (y.shape[1], len(x)), as the reference does."""
    if x.ndim == 1 and y.ndim == 1:
        return np.broadcast_to(x[None, :], (y.size, x.size)).copy(), np.broadcast_to(y[:, None], (y.size, x.size)).copy()
    if x.ndim == 1:
        if x.size != y.shape[1]:
            raise ConfigError(f"grid {name!r}: x and y do not fit")
        return np.broadcast_to(x[None, :], (y.shape[1], x.size)).copy(), y
    if y.ndim == 1:
        if x.shape[0] != y.size:
            raise ConfigError(f"grid {name!r}: x and y do not fit")
        return x, np.broadcast_to(y[:, None], (y.size, x.shape[1])).copy()
    return x, y


def _into_runtime(cfg):
    gamma = cfg.gamma
    try:
        set_gamma(gamma)  # `let _ = GAMMA.set(gamma)`: a second value is ignored (parsing.rs:48)
    except _lib.SbpError as e:
        if e.status != _lib.ERR_GAMMA_SET:
            raise
    grids = dict(cfg.grids)
    default = grids.pop("default", GridConfig())
    names = list(grids)
    index = {n: i for i, n in enumerate(names)}
    kept = []

    # initial conditions (parsing.rs:9-20)
    kind, payload = _one_key(cfg.initial_conditions, "initial_conditions")
    if kind == "vortex":
        ic = _vortex_params(payload)
    elif kind == "expressions":
        ic = _expressions(payload)
        kept.append(ic)
    else:
        raise ConfigError(f"unknown initial conditions {kind!r}")

    # boundary data (parsing.rs:53-69)
    bcfg = cfg.boundary_conditions
    if bcfg == "not_needed":
        bdata = None
    elif bcfg == "initial_conditions":
        bdata = ic
    else:
        k2, p2 = _one_key(bcfg, "boundary_conditions")
        if k2 == "vortex":
            bdata = _vortex_params(p2)
        elif k2 == "expressions":
            bdata = _expressions(p2)
            kept.append(bdata)
        else:
            raise ConfigError(f"unknown boundary conditions {k2!r}")

    coords, ops, bcs = [], [], []
    for name, g in grids.items():
        gx = g.x if g.x is not None else default.x
        gy = g.y if g.y is not None else default.y
        if gx is None or gy is None:
            raise ConfigError(f"grid {name!r} has no coordinates")
        X, Y = _broadcast(_array_form(gx, f"{name}.x"), _array_form(gy, f"{name}.y"), name)
        if X.shape != Y.shape:
            raise ConfigError(f"grid {name!r}: x and y differ in shape")  # Grid::new
        coords.append((X, Y))
        # operators: the grid's, else the default's; the pair is (eta, xi) (parsing.rs:135-168)
        o, do = g.operators or {}, default.operators or {}
        pick = {}
        for axis in ("xi", "eta"):
            nm = o.get(axis) or do.get(axis)
            if nm is None:
                raise ConfigError(f"No {axis} operator found for grid: {name}")
            if nm not in OPERATORS:
                raise ConfigError(f"unknown operator {nm!r}")
            pick[axis] = OPERATORS[nm]()
        ops.append(operators.OperatorPair(pick["eta"], pick["xi"]))
        # boundary conditions: the grid's, else the default's, else the
        # configuration's boundary data (parsing.rs:170-229)
        gb, db = g.boundary_conditions or {}, default.boundary_conditions or {}
        per = {}
        for d in DIRECTIONS:
            b = gb.get(d, db.get(d))
            per[d] = _boundary(b, bdata, index, name)
        bcs.append(Direction(**per))
    return RuntimeConfiguration(names, coords, bcs, ops, cfg.integration_time, ic, gamma, kept)


def _boundary(b, bdata, index, name):
    if b is None:
        if isinstance(bdata, VortexParameters):
            return multigrid.Vortex(bdata)
        if isinstance(bdata, Expressions):
            return multigrid.Eval(bdata)
        raise ConfigError(f"Boundary conditions are not available, but needed for grid {name}")
    if b == "this":
        return multigrid.This()
    if b == "vortex":
        if not isinstance(bdata, VortexParameters):
            raise ConfigError(f"Wanted vortex boundary conditions not found, needed for grid {name}")
        return multigrid.Vortex(bdata)
    kind, p = _one_key(b, f"{name} boundary condition")

    def nb(n):
        if n not in index:
            raise ConfigError(f"grid {name!r}: unknown neighbour {n!r}")
        return index[n]

    if kind in ("neighbour", "neighbor"):
        return multigrid.Grid(nb(p))
    if kind == "interpolate":
        if "operator" not in p or p["operator"] is None:
            raise ConfigError(f"grid {name!r}: interpolation without an operator")
        if p["operator"] not in INTERPOLATION:
            raise ConfigError(f"unknown interpolation operator {p['operator']!r}")
        return multigrid.Interpolate(nb(p.get("neighbour", p.get("neighbor"))), INTERPOLATION[p["operator"]]())
    if kind == "multi":
        return multigrid.MultiGrid([(nb(m.get("neighbour", m.get("neighbor"))), int(m["start"]), int(m["end"])) for m in p])
    raise ConfigError(f"grid {name!r}: unknown boundary condition {kind!r}")


def load(text):
    """Configuration.loads(text).into_runtime()."""
    return Configuration.loads(text).into_runtime()


def run(rt, sys_, steps=None, output=None, number_of_outputs=None):
    """main.rs:96-147: dt = max_dt, ntime = round(T/dt) steps (or `steps`);
    returns (dt, steps taken).

    output: a path for the HDF5 file of the run (hdf5.Output, the layout of
    system.rs:28-117), written as main.rs does: the state at step 0, at the
    end of each stretch of steps_between_outputs = max(number_of_outputs /
    ntime, 1) steps (ntime without number_of_outputs), and at the end. The
    reference's loop takes the next output step as max(itime +
    steps_between_outputs, ntime) (main.rs:123), so its outputs are step 0
    and step ntime, as here."""
    dt = sys_.max_dt()
    sys_.set_dt(dt)
    # f64::round rounds half away from zero
    ntime = int(math.floor(rt.integration_time / dt + 0.5)) if steps is None else steps
    if output is None:
        sys_.advance(ntime)
        return dt, ntime
    from .hdf5 import Output
    if number_of_outputs is not None:
        between = max(number_of_outputs // ntime, 1)  # ntime = 0 fails as n / ntime does in the reference
    else:
        between = ntime
    with Output(output, rt.names, rt.grids) as out:
        out.output(0, sys_.fields())
        itime = 0
        while itime < ntime:
            nexttime = max(itime + between, ntime)
            sys_.advance(nexttime - itime)
            itime = nexttime
            if itime != ntime:
                out.output(itime, sys_.fields())
        out.output(ntime, sys_.fields())
    return dt, ntime
