"""Pins of the oracle (CPU). Each check is one the reference's own tests hold
(polynomial exactness, h2_diff, convergence rates) or one the mathematics
fixes (SBP property, dense brute force, closed-form metrics, free-stream
preservation, vortex translation), chosen so that a dropped term, a wrong
sign or index, or a transposed operand fails.
"""
import numpy as np
import pytest

from common import GAMMA, MACH, VORTEX, cartesian, curvilinear, rel_maxnorm
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


# --- flux, SAT and RK4 pins -------------------------------------------------

def _euler_fluxes(q, gamma=GAMMA):
    """Textbook 2-D Euler fluxes in conservative variables (works on complex
    input for complex-step differentiation)."""
    rho, m, n, e = q
    p = (gamma - 1.0) * (e - (m * m + n * n) / (2.0 * rho))
    E = np.array([m, m * m / rho + p, m * n / rho, m * (e + p) / rho])
    F = np.array([n, m * n / rho, n * n / rho + p, n * (e + p) / rho])
    return E, F


def _flux_jacobian_times(q, d, kx, ky):
    """(kx·∂E/∂q + ky·∂F/∂q)·d at state q by complex-step differentiation
    (exact to rounding: the fluxes are rational in q)."""
    h = 1e-30
    E, F = _euler_fluxes(q.astype(complex) + 1j * h * d)
    return (kx * E.imag + ky * F.imag) / h


def test_fluxes_interior_polynomial_state():
    """RHS_trad with SBP8 on a Cartesian grid with ρ = 1 and momentum /
    energy of degree ≤ 2: every flux is a polynomial of degree ≤ 3, which the
    SBP8 closure and interior differentiate exactly, so away from the SAT
    edges k = -(∂E/∂x + ∂F/∂y) with the textbook fluxes. A dropped pressure
    term, a swapped metric or a wrong sign fails this."""
    nx, ny = 37, 29
    x = np.linspace(-5.0, 5.0, nx)
    y = np.linspace(-2.0, 3.0, ny)
    X, Y = np.meshgrid(x, y)
    q = np.empty((4, ny, nx))
    q[0] = 1.0
    q[1] = 0.5 + 0.1 * X
    q[2] = 0.2 - 0.05 * Y
    q[3] = 3.0 + 0.02 * X * X + 0.01 * X * Y
    S = O.System([(X, Y)], [(O.SBP8, O.SBP8)])
    S.set_rhs_kind(O.RHS_TRAD)
    S.set_state(0, q)
    k = S.rhs(0.0)[0]
    # analytic divergence of the textbook fluxes by complex step along x and y
    h = 1e-30
    dEdx = _euler_fluxes(np.array([q[0], q[1] + 1j * h * 0.1, q[2], q[3] + 1j * h * (0.04 * X + 0.01 * Y)]))[0].imag / h
    dFdy = _euler_fluxes(np.array([q[0], q[1], q[2] + 1j * h * (-0.05), q[3] + 1j * h * (0.01 * X)]))[1].imag / h
    expect = -(dEdx + dFdy)
    inner = (slice(None), slice(1, ny - 1), slice(1, nx - 1))
    scale = np.max(np.abs(expect[inner]))
    assert np.max(np.abs(k[inner] - expect[inner])) < 1e-9 * scale


def _characteristic_penalty(q, d, kx_, ky_, sigma, gamma=GAMMA):
    """Σ_i w_i (λ_i + σ|λ_i|) r_i (l_i · d) over the characteristic fields of
    the normal flux Jacobian A_n = kx_·∂E/∂q + ky_·∂F/∂q at state q, with
    the textbook right eigenvectors r_i (entropy, shear, two acoustic waves;
    unit normal n = k/|k|) and l_i the rows of R⁻¹. The reference's TI
    (euler/src/lib.rs:1213-1220) is diag(1, 1/ρ, 1, 1)·R⁻¹, and its T
    (:1194-1199) is R·diag(1, 1, α, α), α = ρ/(√2 c), except that the
    energy entry of the shear column is ρ(ky·u - kx·v) where the eigenvector
    has (ky·u - kx·v). So the fields carry weights (1, 1/ρ, α, α), the shear
    field feeds the energy equation ρ times too strongly, and T·TI is not the
    identity unless ρ = 1 and α = 1; the reference's SAT is reproduced as it
    is."""
    rho, m, n_, e = q
    u, v = m / rho, n_ / rho
    p = (gamma - 1.0) * (e - (m * m + n_ * n_) / (2.0 * rho))
    c = np.sqrt(gamma * p / rho)
    kn = np.hypot(kx_, ky_)
    kx, ky = kx_ / kn, ky_ / kn
    theta = kx * u + ky * v
    H = (e + p) / rho
    R = np.array([[1.0, 0.0, 1.0, 1.0],
                  [u, ky, u + kx * c, u - kx * c],
                  [v, -kx, v + ky * c, v - ky * c],
                  [(u * u + v * v) / 2.0, ky * u - kx * v, H + c * theta, H - c * theta]])
    U = kx_ * u + ky_ * v
    lam = np.array([U, U, U + c * kn, U - c * kn])
    # the eigenpairs are those of the exact Jacobian (complex step)
    An = np.stack([_flux_jacobian_times(q, np.eye(4)[i], kx_, ky_) for i in range(4)], axis=1)
    assert np.allclose(An @ R, R * lam[None, :], rtol=0, atol=1e-12 * np.max(np.abs(An)))
    alpha = rho / (np.sqrt(2.0) * c)
    w = np.array([1.0, 1.0 / rho, alpha, alpha])
    T = R.copy()
    T[3, 1] *= rho
    return T @ (w * (lam + sigma * np.abs(lam)) * np.linalg.solve(R, d))


@pytest.mark.parametrize("edge", ["north", "south", "east", "west"])
def test_sat_characteristic_supersonic(edge):
    """Characteristic SAT of a uniform supersonic state against far-field
    vortex data g: at the inflow edge every characteristic enters, so the
    penalty is hi·τ·Σ_i w_i·2λ_i·r_i·(l_i·(q-g)) with hi = (n-1)/H[0], and
    at the outflow edge every λ + σ|λ| vanishes. Pins the sign σ, τ, hi and
    the eigenvector tables T and TI of euler/src/lib.rs:1150-1252."""
    nx, ny = 31, 27
    x = np.linspace(-5.0, 5.0, nx)
    y = np.linspace(-2.0, 3.0, ny)
    X, Y = np.meshgrid(x, y)
    speed = 3.0  # sound speed of q is √(1.4/1.3) ≈ 1.04
    vel = {"north": (0.3, -speed), "south": (0.3, speed), "east": (-speed, 0.2), "west": (speed, 0.2)}[edge]
    rho, p = 1.3, 1.0
    qv = np.array([rho, rho * vel[0], rho * vel[1], p / (GAMMA - 1) + 0.5 * rho * (vel[0] ** 2 + vel[1] ** 2)])
    q = np.broadcast_to(qv[:, None, None], (4, ny, nx)).copy()
    d = {"north": O.NORTH, "south": O.SOUTH, "east": O.EAST, "west": O.WEST}[edge]
    S = O.System([(X, Y)], [(O.SBP8, O.UPWIND4)])
    S.set_rhs_kind(O.RHS_TRAD)
    vp = O.VortexParams.make(list(VORTEX), MACH)
    S.set_bc(0, d, O.BC_VORTEX, 0, vp)
    S.set_state(0, q)
    k = S.rhs(0.0)[0]
    g = O.vortex(vp, 0.0, X, Y)  # far-field data at the edge points
    if edge in ("north", "south"):
        n_a, h0, kx, ky = ny, O.op_h(O.SBP8)[0], 0.0, 1.0 / (y[1] - y[0]) / (ny - 1)  # ∇η = (0, dη/dy)
        sl = (slice(None), ny - 1 if edge == "north" else 0, slice(None))
        sigma, tau = (-1.0, 1.0) if edge == "north" else (1.0, -1.0)
        other = (slice(None), 0 if edge == "north" else ny - 1, slice(1, nx - 1))
    else:
        n_a, h0, kx, ky = nx, O.op_h(O.UPWIND4)[0], 1.0 / (x[1] - x[0]) / (nx - 1), 0.0
        sl = (slice(None), slice(None), nx - 1 if edge == "east" else 0)
        sigma, tau = (-1.0, 1.0) if edge == "east" else (1.0, -1.0)
        other = (slice(None), slice(1, ny - 1), 0 if edge == "east" else nx - 1)
    hi = (n_a - 1) / h0
    ge = g[sl]
    expect = np.stack([hi * tau * _characteristic_penalty(qv, qv - ge[:, i], kx, ky, sigma)
                       for i in range(ge.shape[1])], axis=1)
    got = k[sl]
    assert np.max(np.abs(got - expect)) < 1e-11 * np.max(np.abs(expect))
    # the opposite edge is outflow for this state: no penalty there (periodic
    # self-coupling of a uniform state adds none either)
    assert np.max(np.abs(k[other])) < 1e-11 * np.max(np.abs(expect))


def test_rk4_step_is_classical_rk4():
    """One step of the oracle's integrate equals the textbook classical RK4
    composed from its own RHS, with stage times t, t+dt/2, t+dt/2, t+dt (the
    far-field vortex BC makes the RHS time dependent, so a wrong stage time
    fails)."""
    vp = O.VortexParams.make(list(VORTEX), MACH)
    n = 24
    X, Y = cartesian(n, n, -3.0, 3.0)
    S = O.System([(X, Y)], [(O.UPWIND4, O.UPWIND9)])
    for d in range(4):
        S.set_bc(0, d, O.BC_VORTEX, 0, vp)
    S.vortex(0.0, vp)
    t0, dt = 0.3, 0.01
    q0 = S.get_state(0)
    stages = []
    for c, a in ((0.0, 0.0), (0.5, 0.5), (0.5, 0.5), (1.0, 1.0)):
        S.set_state(0, q0 + (a * dt) * stages[-1] if stages else q0)
        stages.append(S.rhs(t0 + c * dt)[0])
    expect = q0 + dt * (stages[0] + 2 * stages[1] + 2 * stages[2] + stages[3]) / 6.0
    S.set_state(0, q0)
    S.time = t0
    S.advance(dt, 1)
    got = S.get_state(0)
    assert rel_maxnorm(got - q0, expect - q0) < 1e-12
    assert abs(S.time - (t0 + dt)) < 1e-15


def _smooth_state(X, Y):
    """A smooth state whose velocities change sign over the grid (so û and v̂
    take both signs) with positive density and pressure."""
    L = np.max(X) - np.min(X)
    s, c = np.sin(2 * np.pi * X / L), np.cos(2 * np.pi * Y / L)
    rho = 1.0 + 0.2 * s * c
    u = 0.8 * np.sin(2 * np.pi * (X + 0.3 * Y) / L)
    v = 0.6 * np.cos(2 * np.pi * (Y - 0.2 * X) / L)
    p = 1.0 + 0.1 * np.cos(2 * np.pi * (X + Y) / L)
    e = p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v)
    return np.array([rho, rho * u, rho * v, e])


def _spectral_radius(q, kx, ky):
    """max |λ| of the normal flux Jacobian kx·∂E/∂q + ky·∂F/∂q at every point
    of q (4, ny, nx): the Jacobian's columns by complex step, its eigenvalues
    by LAPACK."""
    cols = [_flux_jacobian_times(q, np.eye(4)[i][:, None, None], kx, ky) for i in range(4)]
    A = np.moveaxis(np.stack(cols, axis=1), (0, 1), (-2, -1))  # (ny, nx, 4, 4)
    return np.max(np.abs(np.linalg.eigvals(A)), axis=-1)


@pytest.mark.parametrize("ops", [(O.UPWIND9, O.UPWIND9), (O.UPWIND4, O.UPWIND9), (O.UPWIND9H2, O.UPWIND4)])
def test_upwind_dissipation_weight_is_spectral_radius(ops):
    """RHS_upwind - RHS_trad = (DISS_ξ(α_ξ q) + DISS_η(α_η q))/J
    (euler/src/lib.rs:571-616, :619-674), where the dissipation weight of
    each direction is the spectral radius of that direction's normal flux
    Jacobian, J·ξx·∂E/∂q + J·ξy·∂F/∂q and J·ηx·∂E/∂q + J·ηy·∂F/∂q (the
    eigenvalues û, û, û ± c·|J·∇ξ|), derived here from the textbook fluxes by
    complex step and an eigenvalue solver, independently of the oracle's
    formula α = |û| + c·√(J·ξx² + J·ξy²). The state is smooth and non-uniform
    on a curvilinear grid, with velocities of both signs; the SATs (periodic)
    are the same in both right-hand sides and cancel. A halved sound-speed
    term, swapped ξ/η metrics or a dropped absolute value fails."""
    op_eta, op_xi = ops
    ny, nx = 34, 30
    X, Y = curvilinear(nx, ny, amp=0.08)
    q = _smooth_state(X, Y)
    S = O.System([(X, Y)], [(op_eta, op_xi)])
    S.set_state(0, q)
    S.set_rhs_kind(O.RHS_UPWIND)
    k_up = S.rhs(0.0)[0]
    S.set_rhs_kind(O.RHS_TRAD)
    k_trad = S.rhs(0.0)[0]
    J, jxx, jxy, jyx, jyy = S.metrics(0)
    a_xi = _spectral_radius(q, jxx, jxy)
    a_eta = _spectral_radius(q, jyx, jyy)
    expect = np.stack([(O.diff(op_xi, O.DISS, O.XI, a_xi * q[c]) + O.diff(op_eta, O.DISS, O.ETA, a_eta * q[c])) / J
                       for c in range(4)])
    got = k_up - k_trad
    scale = np.max(np.abs(expect))
    assert scale > 1e-3
    assert np.max(np.abs(got - expect)) < 1e-10 * scale
    # both signs of the contravariant velocities occur (a dropped |û| matters)
    uh = jxx * q[1] / q[0] + jxy * q[2] / q[0]
    vh = jyx * q[1] / q[0] + jyy * q[2] / q[0]
    assert np.min(uh) < 0 < np.max(uh) and np.min(vh) < 0 < np.max(vh)


# --- boundary entry points: RHS_no_SAT, SAT functions, explicit tableaus ------

def test_rhs_no_sat_and_sat_functions_compose_the_rhs():
    """RHS_no_SAT (euler/src/lib.rs:518-568) equals the full right-hand side
    away from the block edges, and the four SAT functions applied to it in
    the order of SAT_characteristics (north, south, east, west,
    lib.rs:1002-1013) with the boundary data of boundary_extracts rebuild the
    full right-hand side bit for bit (far-field vortex data on a curvilinear
    upwind grid, so every edge has a penalty)."""
    vp = O.VortexParams.make(list(VORTEX), MACH)
    ny, nx = 31, 36
    X, Y = curvilinear(nx, ny, amp=0.05)
    S = O.System([(X, Y)], [(O.UPWIND4, O.UPWIND9)])
    for d in range(4):
        S.set_bc(0, d, O.BC_VORTEX, 0, vp)
    q = _smooth_state(X, Y)
    S.set_state(0, q)
    t = 0.3
    full = S.rhs(t)[0]
    S.set_sat(False)
    nosat = S.rhs(t)[0]
    S.set_sat(True)
    inner = (slice(None), slice(1, ny - 1), slice(1, nx - 1))
    assert np.array_equal(full[inner], nosat[inner])
    assert np.max(np.abs(full - nosat)) > 1e-3 * np.max(np.abs(full))
    k = nosat
    met = S.metrics(0)
    for d in (O.NORTH, O.SOUTH, O.EAST, O.WEST):
        k = O.sat_edge(O.UPWIND4, O.UPWIND9, d, k, q, met, S.edge(0, d, t))
    assert np.array_equal(k, full)


@pytest.mark.parametrize("edge", ["north", "south", "east", "west"])
def test_sat_function_supersonic_closed_form(edge):
    """One SAT function on arrays (SAT_north / south / east / west,
    lib.rs:1025-1144) against the characteristic sum over exact flux-Jacobian
    eigenvectors, as the system-level test above: uniform supersonic state,
    far-field data from the vortex, inflow edge penalised, nothing else
    touched."""
    nx, ny = 29, 26
    x = np.linspace(-5.0, 5.0, nx)
    y = np.linspace(-2.0, 3.0, ny)
    X, Y = np.meshgrid(x, y)
    speed = 3.0
    vel = {"north": (0.3, -speed), "south": (0.3, speed), "east": (-speed, 0.2), "west": (speed, 0.2)}[edge]
    rho, p = 1.3, 1.0
    qv = np.array([rho, rho * vel[0], rho * vel[1], p / (GAMMA - 1) + 0.5 * rho * (vel[0] ** 2 + vel[1] ** 2)])
    q = np.broadcast_to(qv[:, None, None], (4, ny, nx)).copy()
    d = {"north": O.NORTH, "south": O.SOUTH, "east": O.EAST, "west": O.WEST}[edge]
    met = O.metrics(O.SBP8, O.UPWIND4, X, Y)
    g_full = O.vortex(O.VortexParams.make(list(VORTEX), MACH), 0.0, X, Y)
    if edge in ("north", "south"):
        row = ny - 1 if edge == "north" else 0
        g = g_full[:, row, :]
        sl = (slice(None), row, slice(None))
        n_a, h0, kx, ky = ny, O.op_h(O.SBP8)[0], 0.0, 1.0 / (y[1] - y[0]) / (ny - 1)
        sigma, tau = (-1.0, 1.0) if edge == "north" else (1.0, -1.0)
    else:
        col = nx - 1 if edge == "east" else 0
        g = g_full[:, :, col]
        sl = (slice(None), slice(None), col)
        n_a, h0, kx, ky = nx, O.op_h(O.UPWIND4)[0], 1.0 / (x[1] - x[0]) / (nx - 1), 0.0
        sigma, tau = (-1.0, 1.0) if edge == "east" else (1.0, -1.0)
    k = O.sat_edge(O.SBP8, O.UPWIND4, d, np.zeros_like(q), q, met, g)
    hi = (n_a - 1) / h0
    expect = np.stack([hi * tau * _characteristic_penalty(qv, qv - g[:, i], kx, ky, sigma) for i in range(g.shape[1])],
                      axis=1)
    assert np.max(np.abs(k[sl] - expect)) < 1e-11 * np.max(np.abs(expect))
    rest = np.ones(k.shape, bool)
    rest[sl] = False
    assert not np.any(k[rest])


# integrate/src/lib.rs:30-35, :38-43, :45-56, :138-151
RK4_TAB = ([[0, 0, 0, 0], [0.5, 0, 0, 0], [0, 0.5, 0, 0], [0, 0, 1.0, 0]], [1 / 6, 1 / 3, 1 / 3, 1 / 6], [0.5, 0.5, 1.0])
RK38_TAB = ([[0, 0, 0, 0], [1 / 3, 0, 0, 0], [-1 / 3, 1.0, 0, 0], [1.0, -1.0, 1.0, 0]], [1 / 8, 3 / 8, 3 / 8, 1 / 8],
            [1 / 3, 2 / 3, 1.0])
MID_TAB = ([[0, 0], [0.5, 0]], [0.0, 1.0], [0.5])
EULER_TAB = ([[0.0]], [1.0], [])
BS_TAB = ([[0, 0, 0, 0], [0.5, 0, 0, 0], [0, 0.75, 0, 0], [2 / 9, 1 / 3, 4 / 9, 0]], [2 / 9, 1 / 3, 4 / 9, 0.0],
          [0.0, 0.5, 0.75])


def _farfield_system(n=22):
    vp = O.VortexParams.make(list(VORTEX), MACH)
    X, Y = cartesian(n, n, -3.0, 3.0)
    S = O.System([(X, Y)], [(O.UPWIND4, O.UPWIND9)])
    for d in range(4):
        S.set_bc(0, d, O.BC_VORTEX, 0, vp)
    S.vortex(0.0, vp)
    return S


def test_integrate_tableau_rk4_and_bogacki_shampine_match_fused_steps():
    """The generic tableau step with Rk4's arrays is the oracle's RK4 step
    bit for bit (pinned as classical RK4 above), and with BogackiShampine's
    arrays (C starting with its 0, integrate/src/lib.rs:150) the embedded
    step's solution bit for bit."""
    a, b = _farfield_system(), _farfield_system()
    a.time = b.time = 0.2
    a.advance(0.01, 3)
    b.integrate(*RK4_TAB, 0.01, 3)
    assert np.array_equal(a.get_state(0), b.get_state(0)) and a.time == b.time
    a, b = _farfield_system(), _farfield_system()
    a.embedded_step(0.0, 0.01)
    b.integrate(*BS_TAB, 0.01, 1)
    assert np.array_equal(a.get_state(0), b.get_state(0))


@pytest.mark.parametrize("name", ["rk4_38", "midpoint", "euler"])
def test_integrate_tableau_composed_from_rhs(name):
    """Rk4_38, MidpointMethod and EulerMethod (integrate/src/lib.rs:37-56):
    one step equals the method composed by hand from the oracle's RHS at the
    reference's stage times t + dt·C[i-1] (far-field vortex data, so a wrong
    stage time fails), with the reference's order of scaled_adds."""
    A, B, C = {"rk4_38": RK38_TAB, "midpoint": MID_TAB, "euler": EULER_TAB}[name]
    S = _farfield_system()
    t0, dt = 0.25, 0.01
    q0 = S.get_state(0)
    ks = []
    for i in range(len(B)):
        u = q0.copy()
        for j in range(i):
            if A[i][j] != 0.0:
                u = u + (A[i][j] * dt) * ks[j]
        S.set_state(0, u)
        ks.append(S.rhs(t0 + (dt * C[i - 1] if i else 0.0))[0])
    expect = q0.copy()
    for j in range(len(B)):
        if B[j] != 0.0:
            expect = expect + (B[j] * dt) * ks[j]
    S.set_state(0, q0)
    S.time = t0
    S.integrate(A, B, C, dt, 1)
    assert np.array_equal(S.get_state(0), expect)
    assert S.time == t0 + dt


def test_inner_threads_do_not_change_results():
    """The oracle's threads inside a block (the CPU baseline's all-core
    variant) split rows and points into ranges computed with the same
    arithmetic: results bit for bit those of one thread."""
    X, Y = curvilinear(300, 260, amp=0.07)
    q = _smooth_state(X, Y)
    out = []
    for n in (1, 5):
        O.set_inner_threads(n)
        try:
            S = O.System([(X, Y)], [(O.UPWIND9, O.UPWIND4)])
            S.set_state(0, q)
            S.advance(1e-3, 2)
            out.append(S.get_state(0))
        finally:
            O.set_inner_threads(1)
    assert np.array_equal(out[0], out[1])
