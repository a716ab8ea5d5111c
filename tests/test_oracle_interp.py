"""Pins of the oracle's 2:1 interface interpolation and the Interpolate /
MultiGrid boundary kinds (CPU only).

- the reference's own operator test (test_interpolation_operator,
  sbp/src/operators/interpolation.rs:53-74): 101 coarse points on
  [20, 50] (h2 grids for Interpolation9h2), the linear function is
  reproduced within 1e-11 in both directions;
- constants are reproduced to rounding (every row sums to one): a dropped or
  misplaced tap fails it;
- table shapes and jumps as the reference declares them;
- boundary assembly (euler/src/lib.rs:816-849): Interpolate takes the
  source's opposite edge through the operator, fine to coarse when this
  block has fewer points along ξ; MultiGrid copies the pieces one after
  another, each starting on the last point of the previous piece, and halves
  the first point of every piece but the last (worked by hand from the
  reference's loop).
"""
import numpy as np
import pytest

from common import cartesian
from oracle import oracle as O

IOPS = [(O.INTERP4, False), (O.INTERP8, False), (O.INTERP9, False), (O.INTERP9H2, True)]


def h2linspace(start, end, n):
    """sbp/src/utils.rs:133-140: start, start + h/2, ..., end - h/2, end with h = (end - start)/(n - 2)."""
    h = (end - start) / (n - 2)
    x = np.empty(n)
    x[0], x[-1] = start, end
    x[1:-1] = start + h * (np.arange(1, n - 1) - 0.5)
    return x


def _grids(h2, nc=101):
    if h2:
        return h2linspace(20.0, 50.0, nc), h2linspace(20.0, 50.0, 2 * (nc - 1))
    return np.linspace(20.0, 50.0, nc), np.linspace(20.0, 50.0, 2 * nc - 1)


@pytest.mark.parametrize("iop,h2", IOPS)
def test_interpolation_reproduces_linear(iop, h2):
    xc, xf = _grids(h2)
    np.testing.assert_allclose(O.interpolate(iop, False, xc, xf.size), xf, rtol=0, atol=1e-11)
    np.testing.assert_allclose(O.interpolate(iop, True, xf, xc.size), xc, rtol=0, atol=1e-11)


@pytest.mark.parametrize("iop,h2", IOPS)
@pytest.mark.parametrize("nc", [41, 64])
def test_interpolation_reproduces_constants(iop, h2, nc):
    nf = 2 * (nc - 1) if h2 else 2 * nc - 1
    np.testing.assert_allclose(O.interpolate(iop, False, np.full(nc, 3.25), nf), 3.25, rtol=0, atol=1e-13)
    np.testing.assert_allclose(O.interpolate(iop, True, np.full(nf, 3.25), nc), 3.25, rtol=0, atol=1e-13)


def test_interpolation_table_shapes():
    # (block rows, block cols, diag rows, diag cols, jump) per the reference's declarations
    want = {
        O.INTERP4: (((3, 10), (1, 7), (3, 2)), ((10, 7), (2, 4), (4, 1))),
        O.INTERP8: (((4, 23), (1, 17), (0, 2)), ((23, 16), (2, 9), (8, 1))),
        O.INTERP9: (((6, 22), (1, 15), (5, 2)), ((22, 15), (2, 8), (8, 1))),
        O.INTERP9H2: (((6, 22), (1, 16), (4, 2)), ((22, 15), (2, 8), (8, 1))),
    }
    for iop, dirs in want.items():
        for f2c, (bs, ds, jump) in zip((True, False), dirs):
            block, diag, j = O.interp_tables(iop, f2c)
            assert block.shape == bs and diag.shape == ds and j == jump, (iop, f2c)


def test_interpolation_length_mismatch_is_an_error():
    with pytest.raises(O.OracleError):
        O.interpolate(O.INTERP4, False, np.zeros(10), 20)
    with pytest.raises(O.OracleError):
        O.interpolate(O.INTERP9H2, True, np.zeros(19), 10)


def _coded_state(shape, base):
    ny, nx = shape
    j, i = np.mgrid[0:ny, 0:nx]
    return np.stack([base + 1.0 + 0.01 * i + 0.001 * j, base + 0.5 + 0.02 * i, base - 0.25 + 0.03 * j,
                     base + 2.0 + 0.005 * (i + j)]).astype(np.float64)


@pytest.mark.parametrize("iop", [O.INTERP4, O.INTERP9])
def test_interpolate_boundary_takes_the_opposite_edge(iop):
    """A coarse block (21 x 21) over a fine one (41 x 41): the coarse block's
    south edge is the fine block's north row taken fine to coarse, the fine
    block's north edge the coarse south row taken coarse to fine."""
    coarse, fine = cartesian(21, 21), cartesian(41, 41)
    s = O.System([coarse, fine], [(O.UPWIND4, O.UPWIND4)] * 2)
    s.set_bc_interpolate(0, O.SOUTH, 1, iop)
    s.set_bc_interpolate(1, O.NORTH, 0, iop)
    qc, qf = _coded_state((21, 21), 0.0), _coded_state((41, 41), 5.0)
    s.set_state(0, qc)
    s.set_state(1, qf)
    g = s.edge(0, O.SOUTH)
    for c in range(4):
        np.testing.assert_array_equal(g[c], O.interpolate(iop, True, qf[c, -1, :], 21))
    g = s.edge(1, O.NORTH)
    for c in range(4):
        np.testing.assert_array_equal(g[c], O.interpolate(iop, False, qc[c, 0, :], 41))


def test_interpolate_direction_follows_nx():
    """The reference picks fine to coarse by comparing nx, also on east/west
    edges: two blocks with equal nx and 2:1 ny get coarse to fine from the
    longer edge, whose lengths the operator rejects."""
    a, b = cartesian(11, 21), cartesian(11, 41)
    s = O.System([a, b], [(O.UPWIND4, O.UPWIND4)] * 2)
    s.set_bc_interpolate(0, O.EAST, 1, O.INTERP4)
    s.set_state(0, _coded_state((21, 11), 0.0))
    s.set_state(1, _coded_state((41, 11), 1.0))
    with pytest.raises(O.OracleError):
        s.edge(0, O.EAST)


def test_multigrid_assembly_worked_example():
    """North edge of block 0 (17 points) from the south rows a, b of blocks 1
    and 2 (9 points each): [a0/2, a1..a7, b0..b8]. Three pieces (block 3,
    0..5), (block 4, 2..7), (block 1, 0..9): [c0/2, c1, c2, c3, d2/2, d3,
    d4, d5, a0..a8]."""
    shapes = [(9, 17), (9, 9), (9, 9), (9, 9), (9, 9)]
    grids = [cartesian(nx, ny) for ny, nx in shapes]
    s = O.System(grids, [(O.UPWIND4, O.UPWIND4)] * len(grids))
    states = [_coded_state(sh, 10.0 * k) for k, sh in enumerate(shapes)]
    for k, q in enumerate(states):
        s.set_state(k, q)
    a, b, c, d = (states[k][:, 0, :] for k in (1, 2, 3, 4))
    s.set_bc_multigrid(0, O.NORTH, [(1, 0, 9), (2, 0, 9)])
    want = np.concatenate([a[:, :8], b], axis=1)
    want[:, 0] /= 2
    np.testing.assert_array_equal(s.edge(0, O.NORTH), want)
    s.set_bc_multigrid(0, O.NORTH, [(3, 0, 5), (4, 2, 7), (1, 0, 9)])
    want = np.concatenate([c[:, 0:4], d[:, 2:6], a], axis=1)
    want[:, 0] /= 2
    want[:, 4] /= 2
    np.testing.assert_array_equal(s.edge(0, O.NORTH), want)
    # pieces that do not cover the edge exactly are an error
    s.set_bc_multigrid(0, O.NORTH, [(1, 0, 9), (2, 0, 8)])
    with pytest.raises(O.OracleError):
        s.edge(0, O.NORTH)
