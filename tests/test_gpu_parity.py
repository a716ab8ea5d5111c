"""Parity of the HIP path (through the C ABI) against the oracle.

FP64 tolerances: operator application 1e-13 relative max-norm (the kernels
fuse multiply-adds and sum taps in another order; the values are otherwise
the same arithmetic); states per step 1e-12 relative max-norm per component,
the bar of the FP64 drop-in; right-hand sides relative to their largest
component (their components cancel to very different magnitudes): 1e-12 on
shared metrics, RHS_TOL on each side's own metrics (see below).
"""
import numpy as np
import pytest
import torch

from common import GAMMA, MACH, VORTEX, cartesian, curvilinear, density_noise, rel_global, rel_maxnorm
from oracle import oracle as O

pytestmark = pytest.mark.gpu

OPS = ["upwind4", "upwind4h2", "upwind9", "upwind9h2", "sbp4", "sbp8"]
# System RHS on the library's own metrics: their rounding (~1e-13) differs from
# the oracle's and is amplified by idx·Σ|c| when the fluxes are differentiated
# (~130·3.5 at these sizes); states, not right-hand sides, carry the 1e-12 bar.
RHS_TOL = 5e-12


def _ops_mod():
    from summationbyparts_amd import operators
    return operators


def _vp_o(vortices=VORTEX, mach=MACH):
    return O.VortexParams.make(list(vortices), mach)


def _vp_g(vortices=VORTEX, mach=MACH):
    from summationbyparts_amd.euler import VortexParameters, Vortice
    return VortexParameters([Vortice(*v) for v in vortices], mach)


@pytest.mark.parametrize("name", OPS)
@pytest.mark.parametrize("shape", [(16, 17), (23, 40), (101, 37), (64, 128)])
def test_operator_application(gpu, name, shape):
    ops = _ops_mod()
    op = ops.BY_NAME[name]()
    rng = np.random.default_rng(hash((name, shape)) % 2**32)
    ny, nx = shape
    prev = rng.standard_normal((ny, nx))
    d_prev = torch.from_numpy(prev).cuda()
    kinds = [("diffxi", O.DIFF, O.XI), ("diffeta", O.DIFF, O.ETA)]
    if op.upwind() is not None:
        kinds += [("dissxi", O.DISS, O.XI), ("disseta", O.DISS, O.ETA)]
    for meth, kind, axis in kinds:
        fut = torch.full_like(d_prev, np.nan)
        getattr(op, meth)(d_prev, fut)
        ref = O.diff(O.OPERATORS[name], kind, axis, prev)
        err = np.max(np.abs(fut.cpu().numpy() - ref)) / np.max(np.abs(ref))
        assert err < 1e-13, (meth, err)


def test_operator_pitched_rows(gpu):
    """Row pitch larger than nx (a column slice of a wider array)."""
    ops = _ops_mod()
    op = ops.Upwind9()
    rng = np.random.default_rng(7)
    big = torch.from_numpy(rng.standard_normal((40, 70))).cuda()
    prev = big[:, 5:60]
    fut_big = torch.zeros_like(big)
    fut = fut_big[:, 3:58]
    op.diffeta(prev, fut)
    ref = O.diff(O.UPWIND9, O.DIFF, O.ETA, prev.cpu().numpy())
    assert np.max(np.abs(fut.cpu().numpy() - ref)) / np.max(np.abs(ref)) < 1e-13
    assert float(fut_big[:, :3].abs().max()) == 0.0


@pytest.mark.parametrize("pair", [("upwind9", "upwind9"), ("sbp8", "sbp4"), ("upwind4h2", "upwind9")])
def test_metrics(gpu, pair):
    from summationbyparts_amd import euler
    e, x = O.OPERATORS[pair[0]], O.OPERATORS[pair[1]]
    X, Y = curvilinear(57, 43)
    op = _ops_mod().OperatorPair(_ops_mod().BY_NAME[pair[0]](), _ops_mod().BY_NAME[pair[1]]())
    sys = euler.BlockSystem([(X, Y)], [op])
    got = sys.metrics(0)
    ref = O.metrics(e, x, X, Y)
    for c in range(5):
        assert np.max(np.abs(got[c] - ref[c])) <= 1e-13 * np.max(np.abs(ref[c])), c


def _make_pair(names):
    ops = _ops_mod()
    if isinstance(names, str):
        return ops.BY_NAME[names](), (O.OPERATORS[names], O.OPERATORS[names])
    return ops.OperatorPair(ops.BY_NAME[names[0]](), ops.BY_NAME[names[1]]()), (O.OPERATORS[names[0]],
                                                                               O.OPERATORS[names[1]])


def _c2_state(X, Y, seed=1, block=0):
    q = O.vortex(_vp_o(), 0.0, X, Y)
    q[0] += density_noise(seed, block, *X.shape)
    return q


@pytest.mark.parametrize("names,shape,rhs_kind", [
    ("upwind9", (64, 80), O.RHS_AUTO),
    ("upwind9", (96, 33), O.RHS_TRAD),
    ("upwind4", (40, 70), O.RHS_AUTO),
    ("upwind4h2", (35, 50), O.RHS_AUTO),
    ("upwind9h2", (50, 45), O.RHS_AUTO),
    ("sbp4", (45, 38), O.RHS_AUTO),
    ("sbp8", (60, 44), O.RHS_AUTO),
    (("sbp8", "upwind9"), (52, 61), O.RHS_AUTO),
    (("upwind9", "upwind4"), (52, 61), O.RHS_AUTO),
])
def test_rhs_periodic(gpu, names, shape, rhs_kind):
    from summationbyparts_amd import euler
    op, oids = _make_pair(names)
    ny, nx = shape
    X, Y = curvilinear(nx, ny)
    q = _c2_state(X, Y)
    ref_sys = O.System([(X, Y)], [oids])
    ref_sys.set_rhs_kind(rhs_kind)
    ref_sys.set_state(0, q)
    sys = euler.BlockSystem([(X, Y)], [op])
    sys.set_rhs_kind(rhs_kind)
    sys.set_state(0, q)
    k_ref = ref_sys.rhs(0.0)[0]
    k = sys.rhs(0.0)[0]
    assert rel_global(k, k_ref) < RHS_TOL


N_TILE_SHAPES = 13


@pytest.mark.parametrize("shape_index", list(range(N_TILE_SHAPES)))
def test_rhs_all_tile_shapes(gpu, shape_index):
    """Every kernel configuration, on sizes that leave partial last tiles in
    both axes."""
    from summationbyparts_amd import _lib, euler
    assert _lib.lib().sbp_tile_shape_count() == N_TILE_SHAPES
    for nx, ny in [(131, 77), (70, 40), (33, 97), (80, 64)]:
        X, Y = curvilinear(nx, ny)
        q = _c2_state(X, Y, seed=3)
        ref_sys = O.System([(X, Y)], [(O.UPWIND9, O.UPWIND9)])
        ref_sys.set_state(0, q)
        sys = euler.BlockSystem([(X, Y)], [_ops_mod().Upwind9()])
        sys.set_tile_shape(shape_index)
        sys.set_state(0, q)
        assert rel_global(sys.rhs(0.0)[0], ref_sys.rhs(0.0)[0]) < RHS_TOL, (nx, ny)
        # one fused step (all four stage modes) against the oracle
        dt = 0.2 / (nx - 1)
        sys.advance_steps(dt, 1)
        ref_sys.advance(dt, 1)
        assert rel_maxnorm(sys.get_state(0), ref_sys.get_state(0)) < 1e-12, ("step", nx, ny)


def test_advance_timed_and_tile_index(gpu):
    """The timed advance steps the state bit for bit as the plain advance
    does and returns a positive event time; the tile index names the forced
    shape."""
    from summationbyparts_amd import euler
    nx, ny = 96, 80
    X, Y = curvilinear(nx, ny)
    q = _c2_state(X, Y, seed=2)
    dt = 0.2 / (nx - 1)
    a, b = (euler.BlockSystem([(X, Y)], [_ops_mod().Upwind9()]) for _ in range(2))
    for s in (a, b):
        s.set_tile_shape(8)
        s.set_state(0, q)
    a.advance_steps(dt, 9)
    assert b.advance_timed(dt, 9) > 0.0
    assert np.array_equal(a.get_state(0), b.get_state(0))
    assert b.tile_index() == 8 and b.tile_shape() == (16, 24, 2)


def test_rhs_device_buffers(gpu):
    """euler::RHS_upwind on caller-owned device buffers with explicit boundary data."""
    from summationbyparts_amd import euler
    nx, ny = 50, 41
    X, Y = curvilinear(nx, ny)
    q = _c2_state(X, Y, seed=5)
    met = O.metrics(O.UPWIND9, O.UPWIND9, X, Y)
    # periodic boundary data as boundary_extractor builds it (lib.rs:763-802)
    bn, bs, be, bw = q[:, 0, :], q[:, ny - 1, :], q[:, :, 0], q[:, :, nx - 1]
    ref_sys = O.System([(X, Y)], [(O.UPWIND9, O.UPWIND9)])
    ref_sys.set_state(0, q)
    k_ref = ref_sys.rhs(0.0)[0]
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    k = torch.zeros((4, ny, nx), dtype=torch.float64, device="cuda")
    euler.RHS_upwind(_ops_mod().Upwind9(), k, dev(q), dev(met), [dev(bn), dev(bs), dev(be), dev(bw)])
    assert rel_global(k.cpu().numpy(), k_ref) < 1e-12


def test_c1_sbp8_vortex_farfield(gpu):
    """Config C1: 101² Cartesian, SBP8, vortex far-field SATs, dt = 0.002,
    100 steps. Per-step parity for the first steps, final-state parity, and
    the analytic h2 error to 3 significant digits."""
    from summationbyparts_amd import multigrid
    n = 101
    X, Y = cartesian(n, n)
    vo, vg = _vp_o(), _vp_g()
    ref = O.System([(X, Y)], [(O.SBP8, O.SBP8)])
    sys = multigrid.System([(X, Y)], [_ops_mod().SBP8()])
    for d in range(4):
        ref.set_bc(0, d, O.BC_VORTEX, 0, vo)
        sys.set_boundary(0, d, multigrid.Vortex(vg))
    ref.vortex(0.0, vo)
    sys.vortex(0.0, vg)
    assert rel_maxnorm(sys.get_state(0), ref.get_state(0)) < 1e-14
    dt = 0.2 * min(1.0 / (n - 1), 1.0 / (n - 1))
    sys.set_dt(dt)
    for _ in range(5):
        ref.advance(dt, 1)
        sys.advance(1)
        assert rel_maxnorm(sys.get_state(0), ref.get_state(0)) < 1e-12
    ref.advance(dt, 94)
    sys.advance(94)
    assert rel_maxnorm(sys.get_state(0), ref.get_state(0)) < 1e-11
    # the last step (99 -> 100) from the same state: the oracle re-seeded
    # with the GPU's state and time, so the step is compared at the drop-in
    # bar without the drift of 99 earlier steps
    ref.set_state(0, sys.get_state(0))
    ref.time = sys.time
    ref.advance(dt, 1)
    sys.advance(1)
    assert rel_maxnorm(sys.get_state(0), ref.get_state(0)) < 1e-12
    assert abs(sys.time - ref.time) < 1e-15 and abs(sys.time - 0.2) < 1e-14
    # the analytic error of the oracle's own 100-step run, to 3 digits
    full = O.System([(X, Y)], [(O.SBP8, O.SBP8)])
    for d in range(4):
        full.set_bc(0, d, O.BC_VORTEX, 0, vo)
    full.vortex(0.0, vo)
    full.advance(dt, 100)
    e_ref = full.error_vortex(vo)
    e = sys.error(vg)
    assert float(f"{e:.3g}") == float(f"{e_ref:.3g}"), (e, e_ref)


def test_c2_small_curvilinear_upwind9(gpu):
    """Config C2 at 160x144: curvilinear upwind9, vortex + noise, periodic;
    per-step parity over 10 steps (graph-captured path included)."""
    from summationbyparts_amd import euler
    nx, ny = 160, 144
    X, Y = curvilinear(nx, ny)
    q = _c2_state(X, Y, seed=11)
    ref = O.System([(X, Y)], [(O.UPWIND9, O.UPWIND9)])
    ref.set_state(0, q)
    sys = euler.System(X, Y, _ops_mod().Upwind9())
    sys.set_state(0, q)
    dt = 0.2 / (nx - 1)
    for _ in range(3):
        ref.advance(dt, 1)
        sys.advance_upwind(dt)
        assert rel_maxnorm(sys.field(), ref.get_state(0)) < 1e-12
    ref.advance(dt, 16)
    sys.advance_n(dt, 16)
    assert rel_maxnorm(sys.field(), ref.get_state(0)) < 1e-12


def test_multiblock_grid_interfaces(gpu):
    """2x2 periodic block layout with Grid(j) interfaces of matching edges and
    unequal block sizes (multigrid/examples/quad.json layout)."""
    from summationbyparts_amd import multigrid
    widths, heights = [37, 52], [41, 29]
    xs = np.cumsum([0.0, 10.0, 10.0]) - 10.0
    ys = np.cumsum([0.0, 10.0, 10.0]) - 10.0
    grids, ids = [], []
    for r in range(2):
        for c in range(2):
            nx, ny = widths[c], heights[r]
            x = np.linspace(xs[c], xs[c + 1], nx)
            y = np.linspace(ys[r], ys[r + 1], ny)
            grids.append((np.broadcast_to(x[None, :], (ny, nx)).copy(), np.broadcast_to(y[:, None], (ny, nx)).copy()))
            ids.append((r, c))
    bid = {rc: i for i, rc in enumerate(ids)}
    ref = O.System(grids, [(O.UPWIND4, O.UPWIND4)] * 4)
    sys = multigrid.System(grids, [_ops_mod().Upwind4()] * 4)
    for (r, c), i in bid.items():
        nb = {O.NORTH: bid[((r + 1) % 2, c)], O.SOUTH: bid[((r - 1) % 2, c)],
              O.EAST: bid[(r, (c + 1) % 2)], O.WEST: bid[(r, (c - 1) % 2)]}
        for d, j in nb.items():
            ref.set_bc(i, d, O.BC_GRID, j)
            sys.set_boundary(i, d, multigrid.Grid(j))
    vo, vg = _vp_o(), _vp_g()
    ref.vortex(0.0, vo)
    sys.vortex(0.0, vg)
    dt = 0.2 * min(min(1.0 / (w - 1) for w in widths), min(1.0 / (h - 1) for h in heights))
    sys.set_dt(dt)
    for _ in range(3):
        ref.advance(dt, 1)
        sys.advance(1)
        for i in range(4):
            assert rel_maxnorm(sys.get_state(i), ref.get_state(i)) < 1e-12, i
    assert abs(sys.max_dt() - ref.max_dt()) <= 1e-14 * ref.max_dt()
    e_ref, e = ref.error_vortex(vo), sys.error(vg)
    assert abs(e - e_ref) <= 1e-10 * e_ref


def test_energy_and_maxdt_reductions(gpu):
    from summationbyparts_amd import euler
    nx, ny = 257, 130
    X, Y = curvilinear(nx, ny)
    q = _c2_state(X, Y, seed=2)
    sys = euler.System(X, Y, _ops_mod().Upwind9())
    sys.set_state(0, q)
    ref = O.System([(X, Y)], [(O.UPWIND9, O.UPWIND9)])
    ref.set_state(0, q)
    # exact sum of the H-weighted squares (the oracle's sequential sum, like
    # the reference's, carries ~n·eps rounding; the device sum is compensated)
    import math
    from oracle.oracle import op_h
    h = op_h(O.UPWIND9)
    wx, wy = np.ones(nx), np.ones(ny)
    wx[:8], wx[-8:], wy[:8], wy[-8:] = h, h[::-1], h, h[::-1]
    W = (wy / (ny - 1))[:, None] * (wx / (nx - 1))[None, :]
    exact = math.fsum(np.concatenate([(q[c] * q[c] * W).ravel() for c in range(4)]).tolist())
    assert abs(sys.energy() - exact) <= 1e-14 * exact
    e_ref = O.h2_err(O.UPWIND9, O.UPWIND9, q, np.zeros_like(q)) ** 2
    assert abs(sys.energy() - e_ref) <= 1e-12 * e_ref
    assert abs(sys.max_dt() - ref.max_dt()) <= 1e-15 * ref.max_dt()


def test_c2_full_size(gpu):
    """Config C2 at its full 4096² size.

    RHS: euler::RHS_upwind on device buffers with the oracle's own metrics.
    The metrics are themselves SBP derivatives of the coordinates; two FP64
    implementations round them differently by ~1e-13, and differentiating
    the fluxes again amplifies that by idx·Σ|c| ≈ 1.4e4, so RHS parity is
    checked on shared metrics. The fused step then runs on the library's own
    metrics and the state must agree per step to 1e-12, the drop-in bar.
    Also free-stream preservation of RHS_trad on the curvilinear grid."""
    from summationbyparts_amd import euler
    n = 4096
    X, Y = curvilinear(n, n)
    q = _c2_state(X, Y, seed=1)
    ref = O.System([(X, Y)], [(O.UPWIND9, O.UPWIND9)])
    ref.set_state(0, q)
    k_ref = ref.rhs(0.0)[0]
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    dq = dev(q)
    k = torch.empty_like(dq)
    bounds = [dev(q[:, 0, :]), dev(q[:, n - 1, :]), dev(q[:, :, 0]), dev(q[:, :, n - 1])]
    euler.RHS_upwind(_ops_mod().Upwind9(), k, dq, dev(ref.metrics(0)), bounds)
    assert rel_global(k.cpu().numpy(), k_ref) < 1e-12
    del k, dq, k_ref, bounds
    torch.cuda.empty_cache()
    sys = euler.System(X, Y, _ops_mod().Upwind9())
    sys.set_state(0, q)
    dt = 0.2 / (n - 1)
    ref.advance(dt, 1)
    sys.advance_upwind(dt)
    assert rel_maxnorm(sys.field(), ref.get_state(0)) < 1e-12
    del ref
    u = np.empty_like(q)
    u[0], u[1], u[2], u[3] = 1.0, 1.0, 0.0, 1.0 / (GAMMA * MACH**2) / (GAMMA - 1) + 0.5
    sys.set_state(0, u)
    sys.set_rhs_kind(O.RHS_TRAD)
    k = sys.rhs(0.0)[0]
    scale = n * 10.0 * 3.0 / 100.0  # idx * |J·ξx| * |flux| / J
    assert float(np.max(np.abs(k))) < 1e-10 * scale


def test_error_paths(gpu):
    from summationbyparts_amd import _lib, euler
    # folded grid: det J <= 0 (grid.rs:117)
    X, Y = cartesian(30, 30)
    with pytest.raises(_lib.SbpError) as e:
        euler.System(X[:, ::-1].copy(), Y, _ops_mod().Upwind4())
    assert e.value.status == _lib.ERR_DETJ
    # too small for the closure (algos.rs:120-121)
    X, Y = cartesian(12, 30)
    with pytest.raises(_lib.SbpError) as e:
        euler.System(X, Y, _ops_mod().Upwind9())
    assert e.value.status == _lib.ERR_SHAPE
    # not an upwind operator (lib.rs:597)
    X, Y = cartesian(30, 30)
    sys = euler.System(X, Y, _ops_mod().SBP4())
    with pytest.raises(_lib.SbpError) as e:
        sys.advance_upwind(0.01)
    assert e.value.status == _lib.ERR_NOT_UPWIND
    # negative pressure (lib.rs:642)
    sys = euler.System(X, Y, _ops_mod().Upwind4())
    q = np.ones((4, 30, 30))
    q[3] = -1.0
    sys.set_state(0, q)
    with pytest.raises(_lib.SbpError) as e:
        sys.advance_upwind(0.001)
    assert e.value.status == _lib.ERR_PRESSURE
    # mismatched interface edges (lib.rs:1161-1165)
    from summationbyparts_amd import multigrid
    g0, g1 = cartesian(30, 30), cartesian(31, 30)
    ms = multigrid.System([g0, g1], [_ops_mod().Upwind4()] * 2)
    with pytest.raises(_lib.SbpError) as e:
        ms.set_boundary(0, _lib.NORTH, multigrid.Grid(1))
    assert e.value.status == _lib.ERR_EDGE


def test_c2_full_size_last_step(gpu):
    """Config C2 at 4096²: 199 graph-replayed steps on the GPU, then step
    200 from that state on the GPU and in the oracle (re-seeded with the
    GPU's state), compared at the drop-in bar of 1e-12 relative max-norm
    per component. Checks the fused kernels on a state that has evolved for
    the whole run, not only on the initial one."""
    from summationbyparts_amd import euler
    n = 4096
    X, Y = curvilinear(n, n)
    q = _c2_state(X, Y, seed=1)
    sys = euler.System(X, Y, _ops_mod().Upwind9())
    sys.set_state(0, q)
    dt = 0.2 / (n - 1)
    sys.advance_n(dt, 199)
    q199 = sys.field()
    assert np.all(np.isfinite(q199))
    ref = O.System([(X, Y)], [(O.UPWIND9, O.UPWIND9)])
    ref.set_state(0, q199)
    ref.advance(dt, 1)
    sys.advance_upwind(dt)
    assert rel_maxnorm(sys.field(), ref.get_state(0)) < 1e-12
