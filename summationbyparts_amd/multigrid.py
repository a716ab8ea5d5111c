"""Multi-block systems: multigrid's `System` (multigrid/src/system.rs:368-454)
over the fused HIP path. All blocks of one operator pair advance together in
one stage launch per RK4 stage; interface SATs read the neighbour's edge
straight from its stage-input buffer.

    sys = System(grids=[(x0, y0), (x1, y1)], operators=[Upwind4(), Upwind4()])
    sys.set_boundary(0, NORTH, Grid(1)); sys.set_boundary(1, SOUTH, Grid(0)) ...
    sys.vortex(0.0, params); sys.set_dt(sys.max_dt()); sys.advance(100)
"""
from dataclasses import dataclass, field

from . import _lib
from .euler import BlockSystem, VortexParameters

NORTH, SOUTH, EAST, WEST = _lib.NORTH, _lib.SOUTH, _lib.EAST, _lib.WEST


class This:
    """BoundaryCharacteristic::This (euler/src/lib.rs:738): own opposite edge."""


@dataclass
class Grid:
    """BoundaryCharacteristic::Grid(j) (euler/src/lib.rs:739)."""
    block: int


@dataclass
class Vortex:
    """BoundaryCharacteristic::Vortex(params) (euler/src/lib.rs:740)."""
    params: VortexParameters


@dataclass
class Interpolate:
    """BoundaryCharacteristic::Interpolate(j, operator) (euler/src/lib.rs:743):
    block j's opposite edge through a 2:1 interpolation operator
    (operators.Interpolation4() ... Interpolation9h2())."""
    block: int
    operator: object


@dataclass
class Eval:
    """BoundaryCharacteristic::Eval(evaluator) (euler/src/lib.rs:741): the
    expressions (summationbyparts_amd.expressions.Expressions) evaluated on
    the edge's coordinates at every stage time."""
    expressions: object


@dataclass
class MultiGrid:
    """BoundaryCharacteristic::MultiGrid(pieces) (euler/src/lib.rs:744):
    (block, start, end) pieces of the neighbours' opposite edges."""
    pieces: list = field(default_factory=list)


class System(BlockSystem):
    def __init__(self, grids, operators):
        super().__init__(grids, operators)
        self.dt = float("nan")

    def set_boundary(self, block, direction, bc):
        if isinstance(bc, This) or bc is This:
            self.set_bc(block, direction, _lib.BC_THIS)
        elif isinstance(bc, Grid):
            self.set_bc(block, direction, _lib.BC_GRID, bc.block)
        elif isinstance(bc, Vortex):
            self.set_bc(block, direction, _lib.BC_VORTEX, 0, bc.params)
        elif isinstance(bc, Interpolate):
            self.set_bc_interpolate(block, direction, bc.block, bc.operator.ID)
        elif isinstance(bc, MultiGrid):
            self.set_bc_multigrid(block, direction, bc.pieces)
        elif isinstance(bc, Eval):
            self.set_bc_eval(block, direction, bc.expressions)
        else:
            raise TypeError(f"unsupported boundary condition {bc!r}")

    def set_dt(self, dt):
        self.dt = dt

    def advance(self, nsteps):
        """SingleThreadedSystem::advance (system.rs:491-498)."""
        self.advance_steps(self.dt, nsteps)

    def error(self, params):
        """System::error (system.rs:618-635) against the vortex or the
        expressions of the initial conditions at the system time."""
        if isinstance(params, VortexParameters):
            return self.error_vortex(params)
        return self.error_eval(params)

    def fields(self):
        return [self.get_state(i) for i in range(len(self.shapes))]
