"""Pins of the oracle (CPU). Each check is one the reference's own tests hold
(polynomial exactness, h2_diff, convergence rates) or one the mathematics
fixes (SBP property, dense brute force, closed-form metrics, free-stream
preservation, vortex translation), chosen so that a dropped term, a wrong
sign or index, or a transposed operand fails.
"""
import numpy as np
import pytest

from common import GAMMA, MACH, VORTEX, cartesian, curvilinear
from oracle import oracle as O

ALL = list(O.OPERATORS.items())
UPWIND = [(k, v) for k, v in ALL if k.startswith("upwind")]


def grid_eval(n, f):
    """operators.rs:174-187: x = i/(nx-1), y = j/(ny-1)."""
    ny, nx = n
    x = np.arange(nx) / (nx - 1)
    y = np.arange(ny) / (ny - 1)
    X, Y = np.meshgrid(x, y)
    return f(X, Y)


def check_operator_on(op, n, f, dfdx, dfdy, eps):
    """operators.rs:189-213."""
    x = grid_eval(n, f)
    assert np.max(np.abs(O.diff(op, O.DIFF, O.XI, x) - grid_eval(n, dfdx))) < eps
    assert np.max(np.abs(O.diff(op, O.DIFF, O.ETA, x) - grid_eval(n, dfdy))) < eps


# polynomial checks of upwind9.rs:122-184 / traditional8.rs:66-128 (order <= 4, 1e-4)
POLY8 = [
    (lambda x, y: x + 2 * y, lambda x, y: 1 + 0 * x, lambda x, y: 2 + 0 * x),
    (lambda x, y: x * x + 0.5 * y * y, lambda x, y: 2 * x, lambda x, y: y),
    (lambda x, y: x * y, lambda x, y: y, lambda x, y: x),
    (lambda x, y: x ** 3 + y ** 3 / 6, lambda x, y: 3 * x * x, lambda x, y: y * y / 2),
    (lambda x, y: x * x * y + x * y * y / 2, lambda x, y: 2 * x * y + y * y / 2, lambda x, y: x * x + x * y),
    (lambda x, y: x ** 3 + 2 * x ** 2 * y + 3 * x * y ** 2 + 4 * y ** 3,
     lambda x, y: 3 * x ** 2 + 4 * x * y + 3 * y ** 2, lambda x, y: 2 * x ** 2 + 6 * x * y + 12 * y ** 2),
    (lambda x, y: x ** 4 + x ** 3 * y + x ** 2 * y ** 2 + x * y ** 3 + y ** 4,
     lambda x, y: 4 * x ** 3 + 3 * x ** 2 * y + 2 * x * y ** 2 + y ** 3,
     lambda x, y: x ** 3 + 2 * x ** 2 * y + 3 * x * y ** 2 + 4 * y ** 3),
]


@pytest.mark.parametrize("op", [O.UPWIND9, O.SBP8])
def test_polynomials_order8(op):
    for f, fx, fy in POLY8:
        check_operator_on(op, (16, 32), f, fx, fy, 1e-4)


def test_polynomials_upwind4_sbp4():
    # upwind4.rs:201-240, traditional4.rs:125-156
    for op, n in ((O.UPWIND4, (16, 32)), (O.SBP4, (13, 20))):
        check_operator_on(op, n, POLY8[0][0], POLY8[0][1], POLY8[0][2], 1e-4)
        check_operator_on(op, n, lambda x, y: x * x + 2 * x * y + 3 * y * y, lambda x, y: 2 * x + 2 * y,
                          lambda x, y: 2 * x + 6 * y, 1e-3)
        check_operator_on(op, n, POLY8[5][0], POLY8[5][1], POLY8[5][2], 1e-1)


def h2linspace(start, end, n):
    """sbp/src/utils.rs:133-140."""
    h = (end - start) / (n - 2)
    x = start + h * (np.arange(n) - 0.5)
    x[0], x[-1] = start, end
    return x


def test_h2linspace():
    # utils.rs:142-155
    x = h2linspace(-1.0, 1.0, 50)
    assert abs(x[0] + 1) < 1e-6 and abs(x[49] - 1) < 1e-6
    hend, h = x[1] - x[0], x[2] - x[1]
    assert abs(x[49] - x[48] - hend) < 1e-6 and abs(2 * hend - h) < 1e-6
    assert np.all(np.abs(np.diff(x[1:49]) - h) < 1e-6)


@pytest.mark.parametrize("op", [O.UPWIND4H2, O.UPWIND9H2])
def test_h2_operators_on_h2_grid(op):
    # upwind4h2.rs:105-128 / upwind9h2.rs:100-123: derivative of x, x²/2, x³/3 on the h2 grid
    x = h2linspace(0.0, 1.0, 20)[None, :]
    d = lambda f: O.diff(op, O.DIFF, O.XI, np.ascontiguousarray(f))
    assert np.max(np.abs(d(x) - 1.0)) < 1e-4
    assert np.max(np.abs(d(x * x / 2) - x)) < 1e-4
    assert np.max(np.abs(d(x ** 3 / 3) - x * x)) < 1e-2


def dense(op, kind, n):
    """Dense n x n matrix from the blocks, following sparse_from_block (algos.rs:738-802)."""
    info = O.op_info(op)
    M, N, D = info["M"], info["N"], info["D"]
    start, diag, end = O.op_block(op, kind)
    idx = 1.0 / (1.0 / ((n - 2) if info["is_h2"] else (n - 1)))
    A = np.zeros((n, n))
    A[:M, :N] = start
    for j in range(M, n - M):
        A[j, j - D // 2:j - D // 2 + D] = diag
    A[n - M:, n - N:] = end
    return A * idx


@pytest.mark.parametrize("name,op", ALL)
@pytest.mark.parametrize("n", [16, 17, 23, 40])
def test_dense_brute_force(name, op, n):
    rng = np.random.default_rng(n)
    kinds = [O.DIFF] + ([O.DISS] if O.op_info(op)["is_upwind"] else [])
    info = O.op_info(op)
    if n < max(2 * info["M"], info["N"]):
        pytest.skip("below the operator's minimum size")
    for kind in kinds:
        A = dense(op, kind, n)
        u = rng.standard_normal((5, n))
        got_x = O.diff(op, kind, O.XI, u)
        got_e = O.diff(op, kind, O.ETA, np.ascontiguousarray(u.T))
        ref = u @ A.T
        assert np.max(np.abs(got_x - ref)) < 1e-12 * np.max(np.abs(ref))
        assert np.max(np.abs(got_e.T - ref)) < 1e-12 * np.max(np.abs(ref))


def hnorm(op, n):
    info = O.op_info(op)
    h = O.op_h(op)
    w = np.ones(n)
    w[:info["M"]] = h
    w[n - info["M"]:] = h[::-1]
    return w / ((n - 2) if info["is_h2"] else (n - 1))


@pytest.mark.parametrize("name,op", ALL)
def test_sbp_property(name, op):
    """H·D + (H·D)ᵀ = diag(-1, 0, ..., 0, 1) for the central (DIFF) part, and
    H·DISS symmetric negative semidefinite for the upwind operators."""
    n = 41
    Hm = np.diag(hnorm(op, n))
    D = dense(op, O.DIFF, n)
    Q = Hm @ D
    B = np.zeros((n, n))
    B[0, 0], B[-1, -1] = -1.0, 1.0
    assert np.max(np.abs(Q + Q.T - B)) < 1e-12
    assert np.max(np.abs(D @ np.ones(n))) < 1e-10
    if O.op_info(op)["is_upwind"]:
        S = Hm @ dense(op, O.DISS, n)
        assert np.max(np.abs(S - S.T)) < 1e-12
        assert np.max(np.linalg.eigvalsh((S + S.T) / 2)) < 1e-12
        assert np.max(np.abs(dense(op, O.DISS, n) @ np.ones(n))) < 1e-10


def test_end_block_is_flipped_start():
    for name, op in ALL:
        start, _, end = O.op_block(op, O.DIFF)
        assert np.array_equal(end, -start[::-1, ::-1])
        if O.op_info(op)["is_upwind"]:
            start, _, end = O.op_block(op, O.DISS)
            assert np.array_equal(end, start[::-1, ::-1])


def test_h2_diff():
    # euler/src/lib.rs:434-448
    ones, zeros = np.ones((4, 20, 21)), np.zeros((4, 20, 21))
    for op in (O.UPWIND4, O.UPWIND9, O.SBP4, O.SBP8):
        assert abs(O.h2_err(op, op, ones, zeros) ** 2 - 4.0) < 1e-3


def test_metrics_cartesian_closed_form():
    """On x = a + b·ξ, y = c + d·η the SBP derivatives of linear data are exact."""
    X, Y = cartesian(33, 29, -5.0, 5.0)
    m = O.metrics(O.UPWIND9, O.SBP8, X, Y)
    dx, dy = 10.0, 10.0
    assert np.allclose(m[0], dx * dy, rtol=1e-13)
    assert np.allclose(m[1], dy, rtol=1e-13)
    # derivatives of data constant along an axis: rounding of idx·Σc·5 only
    assert np.max(np.abs(m[2])) < 1e-10 and np.max(np.abs(m[3])) < 1e-10
    assert np.allclose(m[4], dx, rtol=1e-13)


def test_vortex_translation_and_free_stream():
    vp = O.VortexParams.make(list(VORTEX), MACH)
    X, Y = curvilinear(40, 30)
    q1 = O.vortex(vp, 0.7, X, Y)
    q0 = O.vortex(vp, 0.0, X - 0.7, Y)
    assert np.max(np.abs(q1 - q0)) < 1e-12
    far = O.vortex(vp, 0.0, np.full((1, 1), 1e3), np.zeros((1, 1)))[:, 0, 0]
    p_inf = 1.0 / (GAMMA * MACH * MACH)
    assert np.allclose(far, [1.0, 1.0, 0.0, p_inf / (GAMMA - 1) + 0.5], rtol=1e-14)
    # two vortices superpose as deviations from the free stream (vortex.rs:88-111)
    v2 = O.VortexParams.make([VORTEX[0], (2.0, 1.0, 0.4, 0.8)], MACH)
    a = O.vortex(v2, 0.0, X, Y)
    b = O.vortex(O.VortexParams.make([(2.0, 1.0, 0.4, 0.8)], MACH), 0.0, X, Y)
    free = np.array([1.0, 1.0, 0.0, p_inf / (GAMMA - 1) + 0.5])[:, None, None]
    assert np.max(np.abs(a - (O.vortex(vp, 0.0, X, Y) + b - free))) < 1e-12


def test_free_stream_preservation_curvilinear_trad():
    """RHS_trad of a uniform state vanishes on a curvilinear grid (the
    metric terms come from the same SBP operator), and the periodic SATs of a
    uniform state vanish."""
    X, Y = curvilinear(45, 38)
    S = O.System([(X, Y)], [(O.SBP8, O.UPWIND9)])
    S.set_rhs_kind(O.RHS_TRAD)
    u = np.empty((4, 38, 45))
    u[0], u[1], u[2], u[3] = 1.0, 0.7, -0.3, 3.0
    S.set_state(0, u)
    k = S.rhs(0.0)[0]
    assert np.max(np.abs(k)) < 1e-11


def test_rhs_upwind_uniform_cartesian():
    X, Y = cartesian(40, 36)
    S = O.System([(X, Y)], [(O.UPWIND9, O.UPWIND4)])
    u = np.empty((4, 36, 40))
    u[0], u[1], u[2], u[3] = 1.2, 0.5, 0.2, 2.5
    S.set_state(0, u)
    assert np.max(np.abs(S.rhs(0.0)[0])) < 1e-11


def _convergence(op, sizes):
    vp = O.VortexParams.make(list(VORTEX), MACH)
    errs = []
    for n in sizes:
        X, Y = cartesian(n, n)
        S = O.System([(X, Y)], [(op, op)])
        S.vortex(0.0, vp)
        dt = 0.2 * min(1.0 / (n - 1), 1.0 / (n - 1))
        nsteps = int(0.2 / dt)
        S.advance(dt, nsteps)
        ver = O.vortex(vp, nsteps * dt, X, Y)
        errs.append(O.h2_err(op, op, ver, S.get_state(0)))
    qs = [np.log10(errs[i + 1] / errs[i]) / np.log10(sizes[i + 1] / sizes[i]) for i in range(len(sizes) - 1)]
    return errs, qs


@pytest.mark.slow
def test_convergence_upwind4():
    # euler/tests/convergence.rs:54-85: last rate over sizes 25..200 below -4
    errs, qs = _convergence(O.UPWIND4, [25, 35, 50, 71, 100, 150, 200])
    assert qs[-1] < -4.0, qs


@pytest.mark.slow
def test_convergence_upwind9():
    # euler/tests/convergence.rs:87-90: last rate below -8.5
    errs, qs = _convergence(O.UPWIND9, [25, 35, 50, 71, 100, 150, 200])
    assert qs[-1] < -8.5, qs


def test_convergence_short():
    """Cheap version of the convergence check for every operator family."""
    for op, bound in ((O.UPWIND4, -3.5), (O.SBP4, -3.5), (O.UPWIND9, -6.5), (O.SBP8, -6.5)):
        errs, qs = _convergence(op, [50, 71, 100])
        assert qs[-1] < bound, (op, qs)


def test_vortex_farfield_sat_c1_error():
    """Config C1 (101², SBP8, vortex far-field on all edges, 100 steps): the
    analytic error stays at the level the periodic solve reaches, which a
    wrong SAT sign, penalty or edge orientation destroys."""
    vp = O.VortexParams.make(list(VORTEX), MACH)
    n = 101
    X, Y = cartesian(n, n)
    S = O.System([(X, Y)], [(O.SBP8, O.SBP8)])
    for d in range(4):
        S.set_bc(0, d, O.BC_VORTEX, 0, vp)
    S.vortex(0.0, vp)
    S.advance(0.002, 100)
    e = S.error_vortex(vp)
    P = O.System([(X, Y)], [(O.SBP8, O.SBP8)])
    P.vortex(0.0, vp)
    P.advance(0.002, 100)
    e_per = O.h2_err(O.SBP8, O.SBP8, O.vortex(vp, 0.2, X, Y), P.get_state(0))
    assert e < 2.0 * e_per and e < 1e-4, (e, e_per)
    assert abs(S.time - 0.2) < 1e-14


def test_multiblock_equals_single_block():
    """A periodic 2x1 layout glued by Grid interfaces: a uniform state has a
    vanishing RHS, and equal states give equal RHS in both blocks."""
    X, Y = cartesian(30, 30, -5, 5)
    S = O.System([(X, Y), (X + 10.0, Y)], [(O.UPWIND4, O.UPWIND4)] * 2)
    for d, j in ((O.EAST, 1), (O.WEST, 1)):
        S.set_bc(0, d, O.BC_GRID, j)
    for d, j in ((O.EAST, 0), (O.WEST, 0)):
        S.set_bc(1, d, O.BC_GRID, j)
    vp = O.VortexParams.make([], MACH)
    u = np.empty((4, 30, 30))
    u[0], u[1], u[2], u[3] = 1.0, 1.0, 0.0, 1.0 / (GAMMA * MACH ** 2) / (GAMMA - 1) + 0.5
    S.set_state(0, u)
    S.set_state(1, u)
    k = S.rhs(0.0)
    assert max(np.max(np.abs(kk)) for kk in k) < 1e-11
    # identical states with symmetric coupling have the same RHS (the
    # metrics of the shifted copy differ only by rounding of x + 10)
    q = O.vortex(O.VortexParams.make(list(VORTEX), MACH), 0.0, X, Y)
    S.set_state(0, q)
    S.set_state(1, q)
    k = S.rhs(0.0)
    assert np.max(np.abs(k[0] - k[1])) < 1e-12 * np.max(np.abs(k[0]))


def test_max_dt_formula():
    """system.rs:542-589: c_max / max(max|û|·nx, max|v̂|·ny) on a uniform flow."""
    X, Y = cartesian(30, 20)
    S = O.System([(X, Y)], [(O.UPWIND4, O.UPWIND4)])
    u = np.empty((4, 20, 30))
    u[0], u[1], u[2], u[3] = 2.0, 3.0, 1.0, 10.0
    S.set_state(0, u)
    # û = J·ξx·u = 10·1.5, v̂ = J·ηy·v = 10·0.5
    expect = 1.0 / max(15.0 * 30, 5.0 * 20)
    assert abs(S.max_dt() - expect) < 1e-15
