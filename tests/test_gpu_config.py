"""The configuration front end on the GPU: every example configuration of
the reference (multigrid/examples/*.json, kept in tests/golden/configs/)
becomes a system on the device through summationbyparts_amd.config, and is
checked against an oracle system built from the same runtime configuration
(the mapping itself is pinned on the CPU, tests/test_config.py):
  - the initial state (vortex or expressions evaluated on the device, against
    the oracle's evalexpr restatement),
  - max_dt, and three steps of dt = max_dt, the oracle re-seeded with the
    GPU's state before each step (1e-12 relative max-norm per component and
    step over all blocks, the drop-in bar),
and, for single.json and quad.json, the whole run of main.rs (round(T/dt)
steps) against the oracle's own run and its System::error.
Eval boundary data (BoundaryCharacteristic::Eval) comes from the translating
vortex of expressions_vortex.json with its edges left to the configuration's
boundary data, so every edge's data moves with time.

GAMMA is set once per process (a OnceCell, parsing.rs:48 ignores a second
value): this session runs at 1.4, so gresho.json (gamma 1.66666666667) runs
at 1.4 on both sides.
"""
import json
import os

import numpy as np
import pytest

from common import rel_maxnorm_blocks
from oracle import evalexpr as OE
from oracle import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXAMPLES = ["single.json", "double.json", "quad.json", "sedecim.json", "interpolation.json", "mix_operators.json",
            "tsection.json", "expressions_vortex.json", "gresho.json", "smooth_periodic.json"]


def _text(name):
    return open(os.path.join(HERE, "golden", "configs", name)).read()


def _oracle_evaluator(expr):
    form = "conservation" if expr.form == 0 else "pressure"
    return OE.Evaluator(form, expr.globals, expr.exprs, 1.4)


def _oracle_system(rt):
    """The oracle counterpart of a runtime configuration."""
    from summationbyparts_amd import multigrid
    ops = [(O.OPERATORS[type(p.eta).__name__.lower()], O.OPERATORS[type(p.xi).__name__.lower()]) for p in rt.operators]
    ref = O.System(rt.grids, ops)
    for i, bc in enumerate(rt.boundary_conditions):
        for d, b in zip((O.NORTH, O.SOUTH, O.EAST, O.WEST), bc.as_tuple()):
            if isinstance(b, multigrid.This):
                ref.set_bc(i, d, O.BC_THIS)
            elif isinstance(b, multigrid.Grid):
                ref.set_bc(i, d, O.BC_GRID, b.block)
            elif isinstance(b, multigrid.Vortex):
                vp = b.params
                ref.set_bc(i, d, O.BC_VORTEX, 0,
                           O.VortexParams.make([(v.x0, v.y0, v.rstar, v.eps) for v in vp.vortices], vp.mach))
            elif isinstance(b, multigrid.Interpolate):
                ref.set_bc_interpolate(i, d, b.block, b.operator.ID)
            elif isinstance(b, multigrid.MultiGrid):
                ref.set_bc_multigrid(i, d, b.pieces)
            elif isinstance(b, multigrid.Eval):
                ev = _oracle_evaluator(b.expressions)
                ref.set_bc_eval(i, d, lambda t, X, Y, ev=ev: ev.field(t, X, Y))
            else:
                raise AssertionError(b)
    ic = rt.initial_conditions
    if isinstance(ic, multigrid.VortexParameters):
        ref.vortex(0.0, O.VortexParams.make([(v.x0, v.y0, v.rstar, v.eps) for v in ic.vortices], ic.mach))
    else:
        ev = _oracle_evaluator(ic)
        for i, (X, Y) in enumerate(rt.grids):
            ref.set_state(i, ev.field(0.0, X, Y))
    return ref


def _states(s, n):
    return [s.get_state(i) for i in range(n)]


def _check_steps(rt, sys_, ref, steps=3):
    n = len(rt.grids)
    dt = sys_.max_dt()
    assert abs(dt - ref.max_dt()) <= 1e-14 * dt
    sys_.set_dt(dt)
    for k in range(steps):
        before = _states(sys_, n)
        for i in range(n):
            ref.set_state(i, before[i])
        ref.time = sys_.time
        ref.advance(dt, 1)
        sys_.advance(1)
        err = rel_maxnorm_blocks(_states(sys_, n), _states(ref, n))
        assert err < 1e-12, (k, err)
        assert abs(sys_.time - ref.time) <= 1e-15


@pytest.mark.parametrize("name", EXAMPLES)
def test_example_configuration(gpu, name):
    from summationbyparts_amd import config
    rt = config.load(_text(name))
    sys_ = rt.system()
    ref = _oracle_system(rt)
    n = len(rt.grids)
    # initial state: the vortex to rounding, expressions to a few ulps of libm
    assert rel_maxnorm_blocks(_states(sys_, n), _states(ref, n)) < 1e-14
    _check_steps(rt, sys_, ref)


def test_eval_boundary_data(gpu):
    """BoundaryCharacteristic::Eval: the translating vortex of
    expressions_vortex.json as boundary data on all four edges
    ("boundary_conditions": "initial_conditions", no edge set), so the data
    move with the stage times; 5 steps per step against the oracle, then the
    error against the expressions at the system time, and the mixed form
    against FP64."""
    from summationbyparts_amd import _lib, config, multigrid
    d = json.loads(_text("expressions_vortex.json"))
    del d["grids"]["grid"]["boundary_conditions"]
    d["boundary_conditions"] = "initial_conditions"
    rt = config.Configuration.from_value(d).into_runtime()
    assert all(isinstance(b, multigrid.Eval) for b in rt.boundary_conditions[0].as_tuple())
    sys_ = rt.system()
    ref = _oracle_system(rt)
    _check_steps(rt, sys_, ref, steps=5)
    e_gpu = rt.error(sys_)
    X, Y = rt.grids[0]
    ev = _oracle_evaluator(rt.initial_conditions)
    exact = ev.field(sys_.time, X, Y)
    e_ref = O.h2_err(O.UPWIND9, O.UPWIND9, sys_.get_state(0), exact)
    # after 5 steps the error (~4e-9) is a difference of the state and a
    # reference field of size ~1, so the device's and numpy's libm ulps in the
    # reference field bound the agreement: a few ulps of the field's H-norm
    scale = O.h2_err(O.UPWIND9, O.UPWIND9, exact, np.zeros_like(exact))
    assert abs(e_gpu - e_ref) <= 1e-15 * scale, (e_gpu, e_ref)
    # the mixed form on the same Eval edges stays within the FP32 rounding of the RHS
    mixed = rt.system()
    mixed.set_precision(_lib.MIXED)
    fp64 = rt.system()
    dt = fp64.max_dt()
    for s_ in (mixed, fp64):
        s_.set_dt(dt)
        s_.advance(20)
    a, b = mixed.get_state(0), fp64.get_state(0)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-8


@pytest.mark.parametrize("name", ["single.json", "quad.json"])
def test_full_run_and_error(gpu, name):
    """main.rs:96-147: dt = max_dt, round(integration_time/dt) steps, then
    System::error — against the oracle running the same steps."""
    from summationbyparts_amd import config
    rt = config.load(_text(name))
    sys_ = rt.system()
    dt, nsteps = config.run(rt, sys_)
    assert nsteps == int(np.floor(rt.integration_time / dt + 0.5)) and nsteps > 100
    ref = _oracle_system(rt)
    ref.advance(dt, nsteps, 4)
    n = len(rt.grids)
    assert rel_maxnorm_blocks(_states(sys_, n), _states(ref, n)) < 1e-10
    vp = rt.initial_conditions
    e = rt.error(sys_)
    e_ref = ref.error_vortex(O.VortexParams.make([(v.x0, v.y0, v.rstar, v.eps) for v in vp.vortices], vp.mach))
    print(f"{name}: {nsteps} steps of dt {dt:.6g}, error {e:.6e} (oracle {e_ref:.6e})")
    assert abs(e - e_ref) <= 1e-8 * e_ref


def test_run_with_hdf5_output(gpu, tmp_path):
    """main.rs with an output file: the state at step 0 and at step ntime of
    every grid in the reference's layout (system.rs:28-117, 591-616), read
    back by libhdf5 (h5py) where the image has it."""
    from common import have_h5py, read_h5
    from summationbyparts_amd import config
    rt = config.load(_text("quad.json"))
    sys_ = rt.system()
    first = sys_.fields()
    path = tmp_path / "quad.h5"
    dt, nsteps = config.run(rt, sys_, steps=25, output=path)
    last = sys_.fields()
    assert nsteps == 25 and os.path.getsize(path) > 0
    if not have_h5py():
        pytest.skip("no Python with h5py (libhdf5) in this image to read the file back")
    info, arr = read_h5(path, tmp_path)
    assert arr["t"].tolist() == [0, 25]
    assert sorted(info["groups"]) == sorted(rt.names)
    for i, name in enumerate(rt.names):
        assert np.array_equal(arr[f"{name}/x"], rt.grids[i][0]) and np.array_equal(arr[f"{name}/y"], rt.grids[i][1])
        for c, var in enumerate(("rho", "rhou", "rhov", "e")):
            assert np.array_equal(arr[f"{name}/{var}"][0], first[i][c])
            assert np.array_equal(arr[f"{name}/{var}"][1], last[i][c])
    # with number_of_outputs the reference's loop still jumps to ntime
    # (next output step max(itime + steps_between_outputs, ntime), main.rs:123)
    sys2 = rt.system()
    path2 = tmp_path / "quad_n.h5"
    config.run(rt, sys2, steps=10, output=path2, number_of_outputs=5)
    info2, arr2 = read_h5(path2, tmp_path)
    assert arr2["t"].tolist() == [0, 10]


@pytest.mark.parametrize("name", ["interpolation.json", "expressions_vortex.json", "tsection.json"])
@pytest.mark.parametrize("schedule", [0, 1, 2])
def test_example_configuration_each_schedule(gpu, name, schedule):
    """Interpolate, MultiGrid and Eval edges in each stage schedule (split:
    edge tiles beside the interior tiles; merged: one dispatch over all
    tiles; merged after the edge-SAT kernel), three steps per step against
    the oracle."""
    from summationbyparts_amd import config
    rt = config.load(_text(name))
    sys_ = rt.system()
    sys_.set_schedule(schedule)
    assert sys_.schedule() == schedule
    _check_steps(rt, sys_, _oracle_system(rt))
