"""The remaining boundary entry points through the C ABI against the oracle:
RHS_no_SAT, SAT_characteristics and the four SAT functions on caller
buffers, Field::h2_err of two device fields, and integrate over explicit
Butcher tableaus other than the fused RK4.

Tolerances: right-hand sides on shared metrics 1e-12 of their largest
value (the fused kernels sum taps and fuse multiply-adds in another order);
SATs 1e-12 of their largest value; states per step 1e-12 relative max-norm
per component (the drop-in bar); H-norms 1e-13 relative (compensated sums
against an exactly rounded sum).
"""
import math

import numpy as np
import pytest
import torch

from common import MACH, VORTEX, cartesian, curvilinear, density_noise, rel_global, rel_maxnorm
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PAIRS = [("upwind9", "upwind9"), ("upwind4", "upwind9h2"), ("sbp8", "sbp4"), ("upwind4h2", "sbp8")]


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _ops():
    from summationbyparts_amd import operators
    return operators


def _pair(names):
    ops = _ops()
    return ops.OperatorPair(ops.BY_NAME[names[0]](), ops.BY_NAME[names[1]]())


def _state(X, Y, seed):
    q = O.vortex(O.VortexParams.make(list(VORTEX), MACH), 0.0, X, Y)
    q[0] += density_noise(seed, 0, *X.shape)
    return q


@pytest.mark.parametrize("names", PAIRS)
def test_rhs_no_sat_vs_oracle(gpu, names):
    """RHS_no_SAT (euler/src/lib.rs:518-568) with the upwind terms iff both
    operators are upwind, on the oracle's metrics."""
    from summationbyparts_amd import euler
    ny, nx = 45, 52
    X, Y = curvilinear(nx, ny, amp=0.07)
    q = _state(X, Y, 4)
    ref = O.System([(X, Y)], [(O.OPERATORS[names[0]], O.OPERATORS[names[1]])])
    ref.set_state(0, q)
    ref.set_sat(False)
    k_ref = ref.rhs(0.0)[0]
    k = torch.full((4, ny, nx), np.nan, dtype=torch.float64, device="cuda")
    euler.RHS_no_SAT(_pair(names), k, _dev(q), _dev(ref.metrics(0)))
    assert rel_global(k.cpu().numpy(), k_ref) < 1e-12


@pytest.mark.parametrize("names", PAIRS)
def test_sat_functions_vs_oracle(gpu, names):
    """SAT_characteristics on top of RHS_no_SAT rebuilds the full right-hand
    side with far-field vortex data (the oracle's, at t = 0.4), and each SAT
    function alone equals the oracle's on a zero k."""
    from summationbyparts_amd import euler
    ny, nx = 38, 41
    X, Y = curvilinear(nx, ny, amp=0.05)
    q = _state(X, Y, 5)
    oe, ox = O.OPERATORS[names[0]], O.OPERATORS[names[1]]
    vp = O.VortexParams.make(list(VORTEX), MACH)
    ref = O.System([(X, Y)], [(oe, ox)])
    for d in range(4):
        ref.set_bc(0, d, O.BC_VORTEX, 0, vp)
    ref.set_state(0, q)
    t = 0.4
    full = ref.rhs(t)[0]
    met = ref.metrics(0)
    edges = [ref.edge(0, d, t) for d in (O.NORTH, O.SOUTH, O.EAST, O.WEST)]
    op = _pair(names)
    dq, dm = _dev(q), _dev(met)
    k = torch.empty((4, ny, nx), dtype=torch.float64, device="cuda")
    euler.RHS_no_SAT(op, k, dq, dm)
    euler.SAT_characteristics(op, k, dq, dm, [_dev(e) for e in edges])
    assert rel_global(k.cpu().numpy(), full) < 1e-12
    for name, d, e in zip(("north", "south", "east", "west"), (O.NORTH, O.SOUTH, O.EAST, O.WEST), edges):
        k0 = torch.zeros((4, ny, nx), dtype=torch.float64, device="cuda")
        euler.SAT_FUNCTIONS[name](op, k0, dq, dm, _dev(e))
        expect = O.sat_edge(oe, ox, d, np.zeros_like(q), q, met, e)
        got = k0.cpu().numpy()
        assert np.max(np.abs(got - expect)) <= 1e-12 * np.max(np.abs(expect)), name
        assert np.count_nonzero(got) == np.count_nonzero(expect), name


@pytest.mark.parametrize("names,shape", [(("upwind9", "upwind9"), (300, 257)), (("sbp4", "upwind4h2"), (33, 1030)),
                                         (("upwind9h2", "sbp8"), (1500, 40))])
def test_h2_err_two_fields(gpu, names, shape):
    """Field::h2_err (lib.rs:377-432) of two device fields: the compensated
    device sum against the exactly rounded sum of the H-weighted squares and
    against the oracle's sequential sum."""
    from summationbyparts_amd import euler
    ny, nx = shape
    rng = np.random.default_rng(ny * nx)
    a = rng.standard_normal((4, ny, nx))
    b = a + 1e-3 * rng.standard_normal((4, ny, nx))
    oe, ox = O.OPERATORS[names[0]], O.OPERATORS[names[1]]
    got = euler.h2_err(_dev(a), _dev(b), _pair(names))
    he, hx = O.op_h(oe), O.op_h(ox)
    ie, ix = O.op_info(oe), O.op_info(ox)
    wx, wy = np.ones(nx), np.ones(ny)
    wx[:len(hx)], wx[nx - len(hx):] = hx, hx[::-1]
    wy[:len(he)], wy[ny - len(he):] = he, he[::-1]
    W = (wy / ((ny - 2) if ie["is_h2"] else (ny - 1)))[:, None] * (wx / ((nx - 2) if ix["is_h2"] else (nx - 1)))[None, :]
    exact = math.sqrt(math.fsum(((a - b) ** 2 * W[None]).ravel().tolist()))
    assert abs(got - exact) <= 1e-13 * exact
    assert abs(got - O.h2_err(oe, ox, a, b)) <= 1e-12 * exact
    # the reference's own test (h2_diff, lib.rs:434-448): ones against zeros
    ones = np.ones((4, 21, 20))
    assert abs(euler.h2_err(_dev(ones), _dev(np.zeros_like(ones)), _ops().Upwind4()) ** 2 - 4.0) < 1e-3


def test_h2_err_rejects():
    from summationbyparts_amd import _lib, euler
    a = torch.zeros((4, 10, 30), dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.SbpError) as e:
        euler.h2_err(a, a, _ops().Upwind9())  # ny = 10 < 2M = 16
    assert e.value.status == _lib.ERR_SHAPE


def _tableau_systems(grid, names):
    from summationbyparts_amd import multigrid
    from summationbyparts_amd.euler import VortexParameters, Vortice
    vg = VortexParameters([Vortice(*VORTEX[0])], MACH)
    vo = O.VortexParams.make(list(VORTEX), MACH)
    oe, ox = O.OPERATORS[names[0]], O.OPERATORS[names[1]]
    s = multigrid.System([grid], [_pair(names)])
    r = O.System([grid], [(oe, ox)])
    for d in range(4):
        s.set_boundary(0, d, multigrid.Vortex(vg))
        r.set_bc(0, d, O.BC_VORTEX, 0, vo)
    s.vortex(0.0, vg)
    r.vortex(0.0, vo)
    return s, r


@pytest.mark.parametrize("tab", ["Rk4", "Rk4_38", "MidpointMethod", "EulerMethod", "Rk6", "Fehlberg", "BogackiShampine"])
def test_integrate_tableaus_vs_oracle(gpu, tab):
    """integrate::<B> for the reference's tableaus (integrate/src/lib.rs:30-151)
    on a curvilinear block with far-field vortex edges (time-dependent
    boundary data, so the stage times C[i-1] matter): per step against the
    oracle's generic tableau step, re-seeded from the GPU state each step;
    Rk4's arrays also against the oracle's fused-order RK4."""
    from summationbyparts_amd import integrate
    T = getattr(integrate, tab)
    ns, A, B, C = integrate.tableau_arrays(T)
    grid = curvilinear(48, 44, amp=0.06)
    s, r = _tableau_systems(grid, ("upwind9", "upwind4"))
    dt = 0.2 / 47 * (0.25 if tab == "EulerMethod" else 1.0)
    s.time = r.time = 0.1
    for step in range(3):
        r.set_state(0, s.get_state(0))
        r.time = s.time
        r.integrate(A, B, C[:ns - 1], dt, 1)
        s.integrate(T, dt, 1)
        assert rel_maxnorm(s.get_state(0), r.get_state(0)) < 1e-12, step
        assert abs(s.time - r.time) < 1e-15
    if tab == "Rk4":
        q = s.get_state(0)
        r.set_state(0, q)
        r.time = s.time
        r.advance(dt, 1)
        s.integrate(T, dt, 1)
        assert rel_maxnorm(s.get_state(0), r.get_state(0)) < 1e-12


def test_integrate_multiblock_and_rejects(gpu):
    """A generic tableau over two blocks joined by Grid interfaces (the
    neighbour's edge comes from its own stage input), and the mixed form and
    bad tableaus rejected."""
    from summationbyparts_amd import _lib, integrate, multigrid
    ny = 40
    grids = []
    for x0, x1, nx in ((-5.0, 0.0, 36), (0.0, 5.0, 36)):
        x, y = np.linspace(x0, x1, nx), np.linspace(-5.0, 5.0, ny)
        grids.append((np.broadcast_to(x[None, :], (ny, nx)).copy(), np.broadcast_to(y[:, None], (ny, nx)).copy()))
    s = multigrid.System(grids, [_ops().Upwind4()] * 2)
    r = O.System(grids, [(O.UPWIND4, O.UPWIND4)] * 2)
    for i, j in ((0, 1), (1, 0)):
        for d in (_lib.EAST, _lib.WEST):
            s.set_boundary(i, d, multigrid.Grid(j))
            r.set_bc(i, d, O.BC_GRID, j)
    from summationbyparts_amd.euler import VortexParameters, Vortice
    s.vortex(0.0, VortexParameters([Vortice(*VORTEX[0])], MACH))
    r.vortex(0.0, O.VortexParams.make(list(VORTEX), MACH))
    ns, A, B, C = integrate.tableau_arrays(integrate.Rk4_38)
    dt = 0.2 / 35
    for _ in range(2):
        s.integrate(integrate.Rk4_38, dt, 1)
        r.integrate(A, B, C, dt, 1, 2)
    for i in range(2):
        assert rel_maxnorm(s.get_state(i), r.get_state(i)) < 1e-12, i
    s.set_precision(_lib.MIXED)
    with pytest.raises(_lib.SbpError):
        s.integrate(integrate.Rk4_38, dt, 1)
    s.set_precision(_lib.FP64)

    class Big:
        A = tuple(tuple([0.1] * i) for i in range(1, 10))
        B = (0.1,) * 10
        C = (0.1,) * 9
    with pytest.raises(_lib.SbpError):
        s.integrate(Big, dt, 1)


def test_interpolate_bc_rejects_short_edges(gpu):
    """set_bc_interpolate refuses 2:1 couplings whose stencil taps would
    leave the neighbour's edge (the reference's indexing panics there),
    before any launch, and accepts the same operator on long enough edges."""
    from summationbyparts_amd import _lib, multigrid
    # Interpolation9 fine to coarse from a 19-point fine edge onto 10 points:
    # the 23-column block is wider than the fine edge
    nc, nf = 10, 19
    ops = _ops()
    assert _lib.lib().sbp_interpolate(None, _lib.INTERP9, 1, None, nf, None, nc) == _lib.ERR_SHAPE
    g = [cartesian(nc, 20), cartesian(nf, 20)]
    s = multigrid.System(g, [ops.Upwind4()] * 2)
    with pytest.raises(_lib.SbpError) as e:
        s.set_boundary(0, _lib.NORTH, multigrid.Interpolate(1, ops.Interpolation9()))
    assert e.value.status == _lib.ERR_EDGE
    nc, nf = 30, 59
    g = [cartesian(nc, 20), cartesian(nf, 20)]
    s = multigrid.System(g, [ops.Upwind4()] * 2)
    s.set_boundary(0, _lib.NORTH, multigrid.Interpolate(1, ops.Interpolation9()))
    s.set_boundary(1, _lib.SOUTH, multigrid.Interpolate(0, ops.Interpolation9()))
