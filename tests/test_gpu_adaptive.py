"""euler::System::advance_adaptive on the GPU (sbp_system_advance_adaptive)
against the oracle (pinned in tests/test_oracle_adaptive.py).

The accept/reject sequence depends on comparing the H-norm of (previous
state - embedded solution) with maxerr; the cases below keep every
comparison at least 1 % away from a tie (err ≈ 0.435·guess for the periodic
48² block), so the GPU takes the same decisions: the final guess must match
exactly and the state to 1e-12 relative max-norm per component.
"""
import numpy as np
import pytest

from common import MACH, VORTEX, curvilinear, rel_maxnorm
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _vp():
    from summationbyparts_amd.euler import VortexParameters, Vortice
    return VortexParameters([Vortice(*v) for v in VORTEX], MACH), O.VortexParams.make(list(VORTEX), MACH)


def test_advance_adaptive_periodic_vs_oracle(gpu):
    """euler::System semantics: one periodic upwind9 block, RHS_upwind; the
    first two trial steps are rejected, then steps are accepted and grow."""
    from summationbyparts_amd import euler, operators
    n = 48
    X, Y = curvilinear(n, n)
    vg, vo = _vp()
    s = euler.System(X, Y, operators.Upwind9())
    s.vortex(0.0, vg)
    ref = O.System([(X, Y)], [(O.UPWIND9, O.UPWIND9)])
    ref.set_rhs_kind(O.RHS_UPWIND)
    ref.vortex(0.0, vo)
    g = s.advance_adaptive(0.2, 0.04, 0.0125)
    g_ref = ref.advance_adaptive(0.2, 0.04, 0.0125)
    print("final guess", g, g_ref)
    assert g == g_ref
    assert rel_maxnorm(s.field(), ref.get_state(0)) < 1e-12
    # a second call continues from the returned guess
    g = s.advance_adaptive(0.1, g, 0.0125)
    g_ref = ref.advance_adaptive(0.1, g_ref, 0.0125)
    assert g == g_ref
    assert rel_maxnorm(s.field(), ref.get_state(0)) < 1e-12


def test_advance_adaptive_farfield_stage_times(gpu):
    """With far-field vortex edges the boundary data depends on the stage
    times (the loop time plus dt·C[i-1]); every step accepted."""
    from summationbyparts_amd import _lib, euler, operators
    n = 40
    X, Y = curvilinear(n, n)
    vg, vo = _vp()
    s = euler.BlockSystem([(X, Y)], [operators.Upwind9()])
    ref = O.System([(X, Y)], [(O.UPWIND9, O.UPWIND9)])
    for d in range(4):
        s.set_bc(0, d, _lib.BC_VORTEX, 0, vg)
        ref.set_bc(0, d, O.BC_VORTEX, 0, vo)
    s.vortex(0.0, vg)
    ref.vortex(0.0, vo)
    g = s.advance_adaptive_steps(0.05, 0.004, np.inf)
    g_ref = ref.advance_adaptive(0.05, 0.004, np.inf)
    assert g == g_ref
    assert rel_maxnorm(s.get_state(0), ref.get_state(0)) < 1e-12


def test_advance_adaptive_rejects_mixed_and_bad_args(gpu):
    from summationbyparts_amd import _lib, euler, operators
    X, Y = curvilinear(40, 40)
    s = euler.System(X, Y, operators.Upwind9())
    s.set_precision(_lib.MIXED)
    with pytest.raises(_lib.SbpError):
        s.advance_adaptive(0.1, 0.01, 1.0)
    s.set_precision(_lib.FP64)
    with pytest.raises(_lib.SbpError):
        s.advance_adaptive(0.1, -0.01, 1.0)
    t = euler.System(X, Y, operators.SBP8())
    with pytest.raises(_lib.SbpError):
        t.advance_adaptive(0.1, 0.01, 1.0)


def test_advance_adaptive_rejects_block_step_scales(gpu):
    """One trial step drives the whole system, so a block with its own step
    scale (batched independent solves) is rejected before any work."""
    from summationbyparts_amd import _lib, euler, operators
    X, Y = curvilinear(40, 40)
    s = euler.BlockSystem([(X, Y), (X, Y)], [operators.Upwind9()] * 2)
    s.set_rhs_kind(_lib.RHS_UPWIND)
    s.vortex(0.0, _vp()[0])
    q0 = s.get_state(1)
    s.set_dt_scale(1, 0.5)
    with pytest.raises(_lib.SbpError) as e:
        s.advance_adaptive_steps(0.01, 0.002, np.inf)
    assert e.value.status == _lib.ERR_INVALID
    assert np.array_equal(s.get_state(1), q0)
    s.set_dt_scale(1, 1.0)
    s.advance_adaptive_steps(0.01, 0.002, np.inf)
