"""SBP operators: the `SbpOperator2d` / `UpwindOperator2d` surface of
sbp/src/operators.rs:55-150, backed by the HIP kernels of libsbpgpu.

    op = Upwind9()
    op.diffxi(prev, fut)     # sbp/src/operators.rs:56-61
    op.diffeta(prev, fut)    # :62-64
    op.upwind().dissxi(prev, fut)   # :91-96
    (eta_op, xi_op) = OperatorPair(SBP8(), Upwind9())   # tuple impl :115-150

`prev`/`fut` are (ny, nx) float64 arrays: torch tensors on the GPU are used in
place (row pitch = stride(0)); numpy arrays are staged through device memory
and the result is written into `fut`, as the reference writes its output view.
"""
import ctypes

import numpy as np

from . import _lib


def _is_torch(a):
    return type(a).__module__.startswith("torch")


def _device_call(fn, op_id, prev, fut, what):
    import torch

    if prev.shape != fut.shape or prev.dim() != 2:
        raise ValueError(f"{what}: shape mismatch {tuple(prev.shape)} vs {tuple(fut.shape)}")
    for a in (prev, fut):
        if a.dtype != torch.float64 or not a.is_cuda or a.stride(1) != 1:
            raise ValueError(f"{what}: need float64 CUDA tensors with unit column stride")
    ny, nx = prev.shape
    ctx = _lib.context()
    _lib.wait_for_torch()
    st = fn(ctx, op_id, ctypes.c_void_p(prev.data_ptr()), prev.stride(0), ctypes.c_void_p(fut.data_ptr()),
            fut.stride(0), ny, nx)
    _lib.check(st, what)
    _lib.check(_lib.lib().sbp_context_synchronize(ctx), what)


def _apply(fn_name, op_id, prev, fut):
    fn = getattr(_lib.lib(), fn_name)
    if _is_torch(prev):
        _device_call(fn, op_id, prev, fut, fn_name)
        return
    import torch

    p = np.ascontiguousarray(prev, dtype=np.float64)
    if p.ndim == 1:
        raise ValueError("1-D arrays: use a (1, n) or (n, 1) view")
    dp = torch.from_numpy(p).cuda()
    df = torch.empty_like(dp)
    _device_call(fn, op_id, dp, df, fn_name)
    fut[...] = df.cpu().numpy()


class _Operator:
    ID = None
    NAME = None

    def __init__(self):
        M, N, D, h2, up = (ctypes.c_int() for _ in range(5))
        _lib.check(_lib.lib().sbp_operator_info(self.ID, ctypes.byref(M), ctypes.byref(N), ctypes.byref(D),
                                                ctypes.byref(h2), ctypes.byref(up)), "sbp_operator_info")
        self.M, self.N, self.D = M.value, N.value, D.value
        self._h2, self._up = bool(h2.value), bool(up.value)

    # SbpOperator2d
    def diffxi(self, prev, fut):
        _apply("sbp_diffxi", self.ID, prev, fut)

    def diffeta(self, prev, fut):
        _apply("sbp_diffeta", self.ID, prev, fut)

    def h(self):
        out = np.zeros(self.M)
        _lib.check(_lib.lib().sbp_operator_h(self.ID, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                   "sbp_operator_h")
        return out

    def hxi(self):
        return self.h()

    def heta(self):
        return self.h()

    def is_h2(self):
        return self._h2

    def is_h2xi(self):
        return self._h2

    def is_h2eta(self):
        return self._h2

    def op_xi(self):
        return self

    def op_eta(self):
        return self

    def upwind(self):
        return self if self._up else None

    # UpwindOperator2d
    def dissxi(self, prev, fut):
        _apply("sbp_dissxi", self.ID, prev, fut)

    def disseta(self, prev, fut):
        _apply("sbp_disseta", self.ID, prev, fut)

    def coefficients(self, kind=_lib.DIFF):
        """(start, diag, end) blocks as the kernels use them."""
        start, diag, end = np.zeros((self.M, self.N)), np.zeros(self.D), np.zeros((self.M, self.N))
        dp = ctypes.POINTER(ctypes.c_double)
        _lib.check(_lib.lib().sbp_operator_coefficients(self.ID, kind, start.ctypes.data_as(dp),
                                                         diag.ctypes.data_as(dp), end.ctypes.data_as(dp)),
                   "sbp_operator_coefficients")
        return start, diag, end

    @property
    def ids(self):
        """(op_eta, op_xi) ids for the C ABI."""
        return (self.ID, self.ID)

    def __repr__(self):
        return f"{self.NAME}"


class Upwind4(_Operator):
    ID, NAME = _lib.UPWIND4, "Upwind4"


class Upwind4h2(_Operator):
    ID, NAME = _lib.UPWIND4H2, "Upwind4h2"


class Upwind9(_Operator):
    ID, NAME = _lib.UPWIND9, "Upwind9"


class Upwind9h2(_Operator):
    ID, NAME = _lib.UPWIND9H2, "Upwind9h2"


class SBP4(_Operator):
    ID, NAME = _lib.SBP4, "SBP4"


class SBP8(_Operator):
    ID, NAME = _lib.SBP8, "SBP8"


BY_NAME = {"upwind4": Upwind4, "upwind4h2": Upwind4h2, "upwind9": Upwind9, "upwind9h2": Upwind9h2, "sbp4": SBP4,
           "sbp8": SBP8}


class OperatorPair:
    """The tuple operator `(eta, xi)` (sbp/src/operators.rs:115-150): diffxi
    uses `xi`, diffeta uses `eta`; upwind only if both are upwind."""

    def __init__(self, eta, xi):
        self.eta, self.xi = eta, xi

    def diffxi(self, prev, fut):
        self.xi.diffxi(prev, fut)

    def diffeta(self, prev, fut):
        self.eta.diffeta(prev, fut)

    def dissxi(self, prev, fut):
        self.xi.dissxi(prev, fut)

    def disseta(self, prev, fut):
        self.eta.disseta(prev, fut)

    def hxi(self):
        return self.xi.h()

    def heta(self):
        return self.eta.h()

    def is_h2xi(self):
        return self.xi.is_h2()

    def is_h2eta(self):
        return self.eta.is_h2()

    def op_xi(self):
        return self.xi

    def op_eta(self):
        return self.eta

    def upwind(self):
        return self if (self.eta.upwind() is not None and self.xi.upwind() is not None) else None

    @property
    def ids(self):
        return (self.eta.ID, self.xi.ID)

    def __repr__(self):
        return f"({self.eta!r}, {self.xi!r})"


def as_pair_ids(op):
    """(op_eta, op_xi) of an operator or an operator pair."""
    return op.ids


class _Interpolation:
    """InterpolationOperator (sbp/src/operators.rs:108-113): 2:1 interface
    interpolation of a 1-D array. `fine`/`coarse` are float64 CUDA tensors
    (used in place) or numpy arrays (staged through device memory); the
    output is fully overwritten. fine = 2·coarse − 1 points (2·(coarse − 1)
    for Interpolation9h2)."""
    ID = None
    NAME = None

    def fine2coarse(self, fine, coarse):
        self._run(1, fine, coarse)

    def coarse2fine(self, coarse, fine):
        self._run(0, coarse, fine)

    def tables(self, fine2coarse):
        """(block, diag, (jump0, jump1)) of one direction."""
        shape = (ctypes.c_int * 6)()
        _lib.check(_lib.lib().sbp_interp_info(self.ID, int(fine2coarse), shape), "sbp_interp_info")
        br, bc, dr, dc, j0, j1 = list(shape)
        block, diag = np.zeros((br, bc)), np.zeros((dr, dc))
        dp = ctypes.POINTER(ctypes.c_double)
        _lib.check(_lib.lib().sbp_interp_coefficients(self.ID, int(fine2coarse), block.ctypes.data_as(dp),
                                                       diag.ctypes.data_as(dp)), "sbp_interp_coefficients")
        return block, diag, (j0, j1)

    def _run(self, f2c, src, dst):
        import torch

        if _is_torch(src):
            for a in (src, dst):
                if a.dtype != torch.float64 or not a.is_cuda or a.dim() != 1 or a.stride(0) != 1:
                    raise ValueError("interpolation: need contiguous 1-D float64 CUDA tensors")
            ctx = _lib.context()
            _lib.wait_for_torch()
            _lib.check(_lib.lib().sbp_interpolate(ctx, self.ID, f2c, ctypes.c_void_p(src.data_ptr()), src.numel(),
                                                  ctypes.c_void_p(dst.data_ptr()), dst.numel()), "sbp_interpolate")
            _lib.check(_lib.lib().sbp_context_synchronize(ctx), "sbp_interpolate")
            return
        s = torch.from_numpy(np.ascontiguousarray(src, dtype=np.float64)).cuda()
        d = torch.empty(len(dst), dtype=torch.float64, device="cuda")
        self._run(f2c, s, d)
        dst[...] = d.cpu().numpy()

    def __repr__(self):
        return self.NAME


class Interpolation4(_Interpolation):
    ID, NAME = _lib.INTERP4, "Interpolation4"


class Interpolation8(_Interpolation):
    ID, NAME = _lib.INTERP8, "Interpolation8"


class Interpolation9(_Interpolation):
    ID, NAME = _lib.INTERP9, "Interpolation9"


class Interpolation9h2(_Interpolation):
    ID, NAME = _lib.INTERP9H2, "Interpolation9h2"
