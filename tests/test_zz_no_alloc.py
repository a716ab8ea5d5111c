"""sbp_h2_err_noalloc: Field::h2_err (euler/src/lib.rs:377-432) of two device
fields through the library's static device scratch, no allocation. Checked
through the C ABI against sbp_h2_err (the allocating entry point, same
kernels) on two calls in a row, the second with a smaller input, and against
the exactly rounded sum; sizes beyond the maximum are refused on the host.
"""
import ctypes
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _h2(L, fn, ctx, oe, ox, a, b):
    out = ctypes.c_double(np.nan)
    st = fn(ctx, oe, ox, a.shape[1], a.shape[2], ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()),
            ctypes.byref(out))
    return st, out.value


def _exact(a, b, oe, ox):
    ny, nx = a.shape[1:]
    he, hx = O.op_h(oe), O.op_h(ox)
    wx, wy = np.ones(nx), np.ones(ny)
    wx[:len(hx)], wx[nx - len(hx):] = hx, hx[::-1]
    wy[:len(he)], wy[ny - len(he):] = he, he[::-1]
    fe = (ny - 2) if O.op_info(oe)["is_h2"] else (ny - 1)
    fx = (nx - 2) if O.op_info(ox)["is_h2"] else (nx - 1)
    W = (wy / fe)[:, None] * (wx / fx)[None, :]
    return math.sqrt(math.fsum(((a - b) ** 2 * W[None]).ravel().tolist()))


def test_h2_err_noalloc_matches_allocating_entry_point(gpu):
    from summationbyparts_amd import _lib
    L = _lib.lib()
    ctx = _lib.context()
    rng = np.random.default_rng(7)
    # two calls in a row on the shared scratch, the second with a smaller input
    for (ny, nx), (oe, ox) in (((1200, 333), (O.UPWIND9, O.UPWIND4)), ((40, 57), (O.SBP8, O.UPWIND9H2))):
        a = rng.standard_normal((4, ny, nx))
        b = a + 1e-2 * rng.standard_normal((4, ny, nx))
        da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
        torch.cuda.synchronize()
        st0, ref = _h2(L, L.sbp_h2_err, ctx, oe, ox, da, db)
        st1, got = _h2(L, L.sbp_h2_err_noalloc, ctx, oe, ox, da, db)
        assert st0 == _lib.OK and st1 == _lib.OK, (st0, st1)
        assert got == ref, (ny, nx, got, ref)
        exact = _exact(a, b, oe, ox)
        assert abs(got - exact) <= 1e-13 * exact
    # sizes beyond the maximum (2^31 - 1 points) are refused on the host,
    # before the (here absent) buffers are looked at
    assert L.sbp_h2_err_noalloc(ctx, O.UPWIND9, O.UPWIND9, 65536, 65536, None, None, None) == _lib.ERR_SIZE
