"""Pins of the oracle's embedded Bogacki-Shampine step and the adaptive loop
of euler::System::advance_adaptive (CPU only).

- one step of integrate_embedded_rk::<BogackiShampine> equals the tableau
  (utils/integrate/src/lib.rs:138-151) composed from the oracle's own RHS:
  fut = q + dt (2/9 k0 + 1/3 k1 + 4/9 k2), fut2 = q + dt (7/24 k0 + 1/4 k1 +
  1/3 k2 + 1/8 k3). integrate evaluates stage i at t + dt·C[i-1] with
  BogackiShampine's C = [0, 1/2, 3/4, 1], so the stage times are t, t,
  t + dt/2, t + 3dt/4 (not the textbook t, t + dt/2, t + 3dt/4, t + dt);
  far-field vortex edges make the RHS time dependent, so either a textbook
  time or a dropped stage fails;
- the loop's accounting in closed form (euler/src/lib.rs:167-184): when every
  step is accepted the loop time grows by twice the trial step per step and
  the guess by 1.05, so the steps taken and the final guess follow from dt
  alone; when none is (maxerr 0), the state is unchanged, the loop time grows
  by the trial step and the guess shrinks by 0.8 until the time reaches dt.
"""
import numpy as np

from common import MACH, VORTEX, cartesian, rel_maxnorm
from oracle import oracle as O


def _farfield_system(n=24):
    vp = O.VortexParams.make(list(VORTEX), MACH)
    X, Y = cartesian(n, n, -3.0, 3.0)
    S = O.System([(X, Y)], [(O.UPWIND4, O.UPWIND9)])
    for d in range(4):
        S.set_bc(0, d, O.BC_VORTEX, 0, vp)
    S.vortex(0.0, vp)
    return S


def test_bogacki_shampine_step_composition():
    S = _farfield_system()
    t0, dt = 0.3, 0.01
    q0 = S.get_state(0)
    k = []
    for y, t in ((q0, t0), (None, t0), (None, t0 + 0.5 * dt), (None, t0 + 0.75 * dt)):
        if y is None:
            i = len(k)
            a = [[0.5], [0.0, 0.75], [2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0]][i - 1]
            y = q0 + dt * sum(aj * kj for aj, kj in zip(a, k))
        S.set_state(0, y)
        k.append(S.rhs(t)[0])
    fut = q0 + dt * (2.0 / 9.0 * k[0] + 1.0 / 3.0 * k[1] + 4.0 / 9.0 * k[2])
    fut2 = q0 + dt * (7.0 / 24.0 * k[0] + 0.25 * k[1] + 1.0 / 3.0 * k[2] + 0.125 * k[3])
    S.set_state(0, q0)
    t1, got2 = S.embedded_step(t0, dt)
    assert t1 == t0 + dt
    assert rel_maxnorm(S.get_state(0) - q0, fut - q0) < 1e-12
    assert rel_maxnorm(got2[0] - q0, fut2 - q0) < 1e-12


def _periodic_system(n=32):
    vp = O.VortexParams.make(list(VORTEX), MACH)
    X, Y = cartesian(n, n)
    S = O.System([(X, Y)], [(O.UPWIND9, O.UPWIND9)])
    S.set_rhs_kind(O.RHS_UPWIND)
    S.vortex(0.0, vp)
    return S


def test_adaptive_accept_all_closed_form():
    g0, dt = 0.004, 0.05
    # steps until the loop time (2 g per accepted step) reaches dt
    sizes, t, g = [], 0.0, g0
    while t < dt:
        sizes.append(g)
        t += g
        t += g
        g *= 1.05
    S = _periodic_system()
    ref = _periodic_system()
    g_out = S.advance_adaptive(dt, g0, np.inf)
    assert g_out == g
    loop_t = 0.0
    for h in sizes:
        loop_t, _ = ref.embedded_step(loop_t, h)
    assert np.array_equal(S.get_state(0), ref.get_state(0))


def test_adaptive_reject_all_closed_form():
    g0, dt = 0.01, 0.025
    t, g = 0.0, g0
    while t < dt:
        t += g
        g *= 0.8
    assert abs(g - g0 * 0.8 ** 4) < 1e-18
    S = _periodic_system()
    q0 = S.get_state(0)
    g_out = S.advance_adaptive(dt, g0, 0.0)
    assert g_out == g
    assert np.array_equal(S.get_state(0), q0)
