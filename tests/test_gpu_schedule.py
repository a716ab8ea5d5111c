"""The three stage schedules against the oracle: split (edge tiles, which
evaluate their edge points' SATs, on a second stream beside the interior
tiles), merged (one dispatch of the edge instantiation over all tiles, edge
tiles first) and merged after the edge-SAT kernel (the SATs of every block
edge evaluated by a separate kernel, read by the edge tiles). Each step at
the drop-in bar (1e-12 relative max-norm per component) on layouts with
far-field, Grid and periodic edges and unequal operator pairs, and the
schedules against each other: FP64 and mixed results are the same arithmetic
at every point, so they agree to the last bits."""
import numpy as np
import pytest

from common import cartesian, curvilinear, density_noise, rel_global, rel_maxnorm
from oracle import oracle as O

pytestmark = pytest.mark.gpu

VORTEX = [(5.0, 0.0, 1.0, 1.0)]  # on the interface: both blocks carry the flow
MACH = 0.5


def _vp():
    from summationbyparts_amd import euler
    return euler.VortexParameters([euler.Vortice(*v) for v in VORTEX], MACH)


def _layout():
    """A curvilinear 52x61 block beside a Cartesian 47x61 block (61 rows
    each): Grid interfaces east and west, far-field vortex data north and
    south; upwind9 in both directions on the first block, (eta upwind4,
    xi upwind9) on the second."""
    Xa, Ya = curvilinear(52, 61, amp=0.05)
    Xb, Yb = cartesian(47, 61)
    Xb = Xb / 2.0 + 7.5  # [5, 10] beside [-5, 5]
    return [(Xa, Ya), (Xb, Yb)]


def _systems(schedule, precision="fp64"):
    from summationbyparts_amd import _lib, multigrid, operators
    grids = _layout()
    sys_ = multigrid.System(grids, [operators.Upwind9(), operators.OperatorPair(operators.Upwind4(), operators.Upwind9())])
    sys_.set_schedule(schedule)
    if precision == "mixed":
        sys_.set_precision(_lib.MIXED)
    ref = O.System(grids, [(O.UPWIND9, O.UPWIND9), (O.UPWIND4, O.UPWIND9)])
    vp = _vp()
    vo = O.VortexParams.make(VORTEX, MACH)
    for b, o in ((0, 1), (1, 0)):
        sys_.set_boundary(b, multigrid.EAST, multigrid.Grid(o))
        sys_.set_boundary(b, multigrid.WEST, multigrid.Grid(o))
        ref.set_bc(b, O.EAST, O.BC_GRID, o)
        ref.set_bc(b, O.WEST, O.BC_GRID, o)
        for d, od in ((multigrid.NORTH, O.NORTH), (multigrid.SOUTH, O.SOUTH)):
            sys_.set_boundary(b, d, multigrid.Vortex(vp))
            ref.set_bc(b, od, O.BC_VORTEX, 0, vo)
    sys_.vortex(0.0, vp)
    ref.vortex(0.0, vo)
    for i, (X, _) in enumerate(grids):
        q = sys_.get_state(i)
        q[0] += density_noise(7, i, *X.shape)
        sys_.set_state(i, q)
        ref.set_state(i, q)
    return sys_, ref


@pytest.mark.parametrize("schedule", [0, 1, 2])
def test_schedule_against_oracle(gpu, schedule):
    sys_, ref = _systems(schedule)
    assert sys_.schedule() == schedule
    dt = 0.2 * 10.0 / 60.0 / 4.0
    sys_.set_dt(dt)
    for _ in range(3):
        for i in range(2):
            ref.set_state(i, sys_.get_state(i))
        ref.time = sys_.time
        ref.advance(dt, 1)
        sys_.advance(1)
        for i in range(2):
            assert rel_maxnorm(sys_.get_state(i), ref.get_state(i)) < 1e-12, (schedule, i)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_schedules_agree(gpu, precision):
    """Ten steps in each schedule at one tile shape (the mixed form's
    rounding follows the tiles' reference states, so the autotuner could
    otherwise pick shapes that round differently)."""
    out = []
    for schedule in (0, 1, 2):
        s, _ = _systems(schedule, precision)
        s.set_tile_shape(0)
        s.set_dt(0.2 * 10.0 / 60.0 / 4.0)
        s.advance(10)
        out.append([s.get_state(i) for i in range(2)])
        s.close()
    for other in out[1:]:
        for a, b in zip(out[0], other):
            assert rel_maxnorm(a, b) < 1e-14


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tile_orders_agree(gpu, precision):
    """Both tile orders (XCD bands, row-major) at one shape and schedule: the
    same tiles run the same arithmetic, only the workgroup order differs, so
    the results agree bit for bit."""
    out = []
    for order in (0, 1):
        s, _ = _systems(0, precision)
        s.set_tile_shape(0)
        s.set_tile_order(order)
        assert s.tile_order() == order
        s.set_dt(0.2 * 10.0 / 60.0 / 4.0)
        s.advance(10)
        out.append([s.get_state(i) for i in range(2)])
        s.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_schedule_c2_block_and_autotune(gpu):
    """A periodic 300x300 curvilinear upwind9 block in each schedule and
    autotuned (one RHS each against the oracle); the autotuned system
    reports the schedule it picked."""
    from summationbyparts_amd import _lib, euler, operators
    n = 300
    X, Y = curvilinear(n, n)
    ks = []
    for schedule in (0, 1, 2, -1):
        s = euler.System(X, Y, operators.Upwind9())
        s.set_schedule(schedule)
        s.vortex(0.0, _vp())
        q = s.field()
        q[0] += density_noise(1, 0, n, n)
        s.set_state(0, q)
        s.set_rhs_kind(_lib.RHS_UPWIND)
        ks.append(s.rhs(0.0)[0])
        assert s.schedule() in (0, 1, 2) and (schedule < 0 or s.schedule() == schedule)
        s.close()
    ref = O.System([(X, Y)], [(O.UPWIND9, O.UPWIND9)])
    ref.set_rhs_kind(O.RHS_UPWIND)
    ref.set_state(0, q)
    k_ref = ref.rhs(0.0)[0]
    # the RHS bar of the parity tests (max-norm over all components: the
    # RHS components cancel to very different magnitudes)
    for k in ks:
        assert rel_global(k, k_ref) < 5e-12
    assert rel_global(ks[0], ks[1]) < 1e-15 and rel_global(ks[0], ks[2]) < 1e-15


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("schedule", [0, 1, 2])
def test_time_update_in_stage_dispatches(gpu, schedule, precision):
    """The RK4 step's time update is a store by one thread of the first two
    stages' dispatches (stage S0 copies the start time into a second slot,
    which stages S1-S3 read; S1 writes that slot + dt into the system time).
    With two operator groups, after single steps and after multi-step calls
    (graphs of several steps), the device time equals the host's running sum
    t += dt (integrate::<Rk4>, utils/integrate/src/lib.rs:176-222) bit for
    bit, and the far-field data follow it: the state equals that of a system
    stepped once per call. Both systems run one tile shape: the mixed form's
    per-tile reference terms make its rounding depend on the tile, which the
    autotuner picks by timing."""
    sys_, _ = _systems(schedule, precision)
    sys_.set_tile_shape(0)
    dt = 0.2 * 10.0 / 60.0 / 4.0
    sys_.set_dt(dt)
    sys_.time = 0.3
    t = 0.3
    for n in (1, 3, 16, 5):
        sys_.advance(n)
        for _ in range(n):
            t += dt
        assert sys_.time == t, (n, sys_.time, t)
    one, _ = _systems(schedule, precision)
    one.set_tile_shape(0)
    one.set_dt(dt)
    one.time = 0.3
    for _ in range(25):
        one.advance(1)
    assert one.time == t
    for i in range(2):
        assert np.array_equal(one.get_state(i), sys_.get_state(i)), i


def test_fp64_tile_shapes_agree(gpu):
    """The FP64 form evaluates every point's face terms and stencil sums in
    the same order whatever the tile (closure rows from the table path and
    interior rows from the window path give the same products): all 13 tile
    shapes step a curvilinear upwind9 block with far-field edges to the same
    bits."""
    from summationbyparts_amd import euler, multigrid, operators
    X, Y = curvilinear(101, 77, amp=0.05)
    vp = _vp()
    out = []
    for k in range(13):
        s = multigrid.System([(X, Y)], [operators.Upwind9()])
        for d in range(4):
            s.set_boundary(0, d, multigrid.Vortex(vp))
        s.vortex(0.0, vp)
        q = s.get_state(0)
        q[0] += density_noise(3, 0, *X.shape)
        s.set_state(0, q)
        s.set_tile_shape(k)
        s.set_dt(0.2 / 100.0)
        s.advance(4)
        assert s.tile_index() == k
        out.append(s.get_state(0))
    for k in range(1, 13):
        assert np.array_equal(out[0], out[k]), k
