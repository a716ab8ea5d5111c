"""CPU checks of the C-ABI library: it loads, exports every function that
include/sbpgpu.h declares, and its host-side tables and traits agree with the
oracle. No compute call is made here (no GPU in this process)."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "sbpgpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(sbp_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_header():
    from summationbyparts_amd import _lib
    L = _lib.lib()
    names = header_functions()
    assert len(names) >= 35
    for n in names:
        assert hasattr(L, n), n
    assert sorted(_lib.EXPORTED) == names


@pytest.mark.parametrize("name,op", list(O.OPERATORS.items()))
def test_tables_match_oracle(name, op):
    from summationbyparts_amd import operators
    g = operators.BY_NAME[name]()
    info = O.op_info(op)
    assert (g.M, g.N, g.D, g.is_h2(), g.upwind() is not None) == (info["M"], info["N"], info["D"], info["is_h2"],
                                                                 info["is_upwind"])
    assert np.array_equal(g.h(), O.op_h(op))
    kinds = [O.DIFF] + ([O.DISS] if info["is_upwind"] else [])
    for kind in kinds:
        for a, b in zip(g.coefficients(kind), O.op_block(op, kind)):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("cls,iop", [("Interpolation4", O.INTERP4), ("Interpolation8", O.INTERP8),
                                     ("Interpolation9", O.INTERP9), ("Interpolation9h2", O.INTERP9H2)])
def test_interpolation_tables_match_oracle(cls, iop):
    """The library's 2:1 interpolation tables equal the oracle's bit for bit
    (both carry the reference's literals)."""
    from summationbyparts_amd import operators
    g = getattr(operators, cls)()
    for f2c in (True, False):
        gb, gd, gj = g.tables(f2c)
        ob, od, oj = O.interp_tables(iop, f2c)
        assert gj == oj and np.array_equal(gb, ob) and np.array_equal(gd, od), f2c


def test_status_strings_and_invalid_args():
    from summationbyparts_amd import _lib
    L = _lib.lib()
    assert _lib.status_string(_lib.ERR_NOT_UPWIND) == "This is not an upwind operator"
    assert _lib.status_string(_lib.ERR_GAMMA) == "GAMMA is not defined"
    M = ctypes.c_int()
    assert L.sbp_operator_info(17, ctypes.byref(M), None, None, None, None) == _lib.ERR_INVALID
    buf = np.zeros(200)
    dp = buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert L.sbp_operator_coefficients(_lib.SBP8, _lib.DISS, dp, dp, dp) == _lib.ERR_NOT_UPWIND
    assert L.sbp_system_advance(None, 0.1, 1) == _lib.ERR_INVALID
    assert L.sbp_tile_shape_count() >= 1


def test_gamma_oncecell_semantics():
    from summationbyparts_amd import _lib
    L = _lib.lib()
    g = ctypes.c_double()
    # this process may not have set gamma yet
    st = L.sbp_get_gamma(ctypes.byref(g))
    assert st in (_lib.OK, _lib.ERR_GAMMA)
    assert L.sbp_set_gamma(1.4) == _lib.OK
    assert L.sbp_set_gamma(1.4) == _lib.OK
    assert L.sbp_set_gamma(1.3) == _lib.ERR_GAMMA_SET
    assert L.sbp_get_gamma(ctypes.byref(g)) == _lib.OK and g.value == 1.4


def test_product_does_not_touch_oracle():
    """The package never imports or loads anything from oracle/."""
    pkg = os.path.join(ROOT, "summationbyparts_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hpp", ".hip", ".h")) or f == "Makefile":
                text = open(os.path.join(dirpath, f), errors="ignore").read()
                assert "oracle" not in text.lower(), os.path.join(dirpath, f)


def _taps_inside(block, diag, jump, nin, nout):
    """Whether every input index the reference's interpolation loop
    (sbp/src/operators/interpolation.rs:26-47) reads for nout outputs lies in
    [0, nin): the loop itself, walked row by row."""
    brows, bcols = block.shape
    drows, dcols = diag.shape
    for i in range(nout):
        if i < brows:
            idx = range(bcols)
        elif nout - i <= brows:
            idx = range(nin - bcols, nin)
        else:
            n = i - brows
            base = jump[0] + jump[1] * (n // drows)
            idx = range(base, base + dcols)
        if idx.start < 0 or idx.stop > nin:
            return False
    return True


@pytest.mark.parametrize("iop", [O.INTERP4, O.INTERP8, O.INTERP9, O.INTERP9H2])
def test_interpolation_rejects_short_edges(iop):
    """sbp_interpolate (checked before its handle, so no device is needed)
    and the oracle reject every length pair whose taps would leave the input,
    and accept every other pair the operator couples. Short coarse edges of
    the wide operators (a 23-column block) are the cases that pass the
    2:1 length relation but not the taps."""
    from summationbyparts_amd import _lib
    L = _lib.lib()
    h2 = iop == O.INTERP9H2
    rejected = 0
    for coarse in range(2, 60):
        fine = 2 * (coarse - 1) if h2 else 2 * coarse - 1
        for f2c in (True, False):
            block, diag, jump = O.interp_tables(iop, f2c)
            nin, nout = (fine, coarse) if f2c else (coarse, fine)
            ok = _taps_inside(block, diag, jump, nin, nout)
            st = L.sbp_interpolate(None, iop, int(f2c), None, nin, None, nout)
            assert st == (_lib.ERR_INVALID if ok else _lib.ERR_SHAPE), (coarse, f2c, st)
            if ok:
                out = O.interpolate(iop, f2c, np.linspace(0.0, 1.0, nin), nout)
                assert out.shape == (nout,)
            else:
                rejected += 1
                with pytest.raises(O.OracleError):
                    O.interpolate(iop, f2c, np.zeros(nin), nout)
        # lengths the operator does not couple are rejected either way
        assert L.sbp_interpolate(None, iop, 1, None, fine + 1, None, coarse) == _lib.ERR_SHAPE
    assert rejected > 0


def test_h2linspace_host():
    """sbp_h2linspace (a host function) against the reference's test
    (sbp/src/utils.rs:142-155) and its formula written out (two roundings
    per point, as Rust computes start + h·(i - 0.5))."""
    from summationbyparts_amd import utils
    x = utils.h2linspace(-1.0, 1.0, 50)
    assert abs(x[0] + 1) < 1e-6 and abs(x[49] - 1) < 1e-6
    hend, h = x[1] - x[0], x[2] - x[1]
    assert abs(x[49] - x[48] - hend) < 1e-6 and abs(2 * hend - h) < 1e-6
    assert np.all(np.abs(np.diff(x[1:49]) - h) < 1e-6)
    for start, end, n in ((-1.0, 1.0, 50), (0.0, 5.0, 50), (-5.0, 0.0, 98), (0.3, 2.7, 2), (1.0, 1.5, 7)):
        got = utils.h2linspace(start, end, n)
        hh = (end - start) / (n - 2) if n > 2 else 0.0
        expect = [start] + [start + hh * (i - 0.5) for i in range(1, n - 1)] + [end]
        assert got.tolist() == expect, (start, end, n)
    from summationbyparts_amd import _lib
    with pytest.raises(_lib.SbpError):
        utils.h2linspace(0.0, 1.0, 1)


def test_h2_err_size_checks_before_handles():
    """sbp_h2_err and sbp_h2_err_noalloc check their sizes on the host before
    the handle: beyond 2^31 - 1 points SBP_ERR_SIZE, below the closures
    SBP_ERR_SHAPE, and only then a missing context (SBP_ERR_INVALID)."""
    from summationbyparts_amd import _lib
    L = _lib.lib()
    for fn in (L.sbp_h2_err, L.sbp_h2_err_noalloc):
        assert fn(None, _lib.UPWIND9, _lib.UPWIND9, 65536, 65536, None, None, None) == _lib.ERR_SIZE
        assert fn(None, _lib.UPWIND9, _lib.UPWIND9, 10, 40, None, None, None) == _lib.ERR_SHAPE
        assert fn(None, _lib.UPWIND9, _lib.UPWIND9, 40, 40, None, None, None) == _lib.ERR_INVALID
        assert fn(None, 17, _lib.UPWIND9, 40, 40, None, None, None) == _lib.ERR_INVALID


def test_system_create_size_checks_before_handles():
    """sbp_system_create checks every block's size on the host before the
    context: SBP_ERR_SHAPE below the operators' closures, SBP_ERR_SIZE when
    the 5 planes of a block (pitch x ny doubles) reach 2^32 bytes (the stage
    kernels' 32-bit offsets), then SBP_ERR_INVALID for the missing context."""
    import ctypes
    from summationbyparts_amd import _lib
    L = _lib.lib()

    def create(shapes, ops=(_lib.UPWIND9, _lib.UPWIND9)):
        n = len(shapes)
        ny = (ctypes.c_int * n)(*[s[0] for s in shapes])
        nx = (ctypes.c_int * n)(*[s[1] for s in shapes])
        oe = (ctypes.c_int * n)(*[ops[0]] * n)
        ox = (ctypes.c_int * n)(*[ops[1]] * n)
        return L.sbp_system_create(None, n, ny, nx, oe, ox, None, None, None)

    # 10 368^2: pitch 10 368, 5 planes = 4.30e9 bytes > 2^32; 10 352^2 fits
    assert create([(64, 64), (10368, 10368)]) == _lib.ERR_SIZE
    assert create([(64, 64), (10352, 10352)]) == _lib.ERR_INVALID
    assert create([(64, 64), (10, 40)]) == _lib.ERR_SHAPE
    assert create([(64, 64)]) == _lib.ERR_INVALID
    assert create([]) == _lib.ERR_INVALID
