"""ctypes binding of libsbpgpu.so (the C ABI declared in include/sbpgpu.h).

The HIP library is the only compute path: when it is missing or fails to load,
every entry point raises instead of falling back to anything else.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsbpgpu.so")

# sbp_status
OK, ERR_INVALID, ERR_SHAPE, ERR_DETJ, ERR_RHO, ERR_PRESSURE, ERR_GAMMA, ERR_EDGE, ERR_HIP, ERR_NOT_UPWIND, ERR_OOM, \
    ERR_GAMMA_SET, ERR_SIZE, ERR_EXPR = range(14)
# sbp_operator
UPWIND4, UPWIND4H2, UPWIND9, UPWIND9H2, SBP4, SBP8 = range(6)
DIFF, DISS = 0, 1
NORTH, SOUTH, EAST, WEST = 0, 1, 2, 3
BC_THIS, BC_GRID, BC_VORTEX, BC_INTERPOLATE, BC_MULTIGRID, BC_EVAL = 0, 1, 2, 3, 4, 5
EXPR_CONSERVATION, EXPR_PRESSURE = 0, 1
INTERP4, INTERP8, INTERP9, INTERP9H2 = 0, 1, 2, 3
RHS_AUTO, RHS_TRAD, RHS_UPWIND = 0, 1, 2
FP64, MIXED = 0, 1


class SbpError(RuntimeError):
    """A failed library call; `status` is the sbp_status code."""

    def __init__(self, status, what):
        self.status = status
        super().__init__(f"{what}: {status_string(status)} (status {status})")


class Vortice(ctypes.Structure):
    _fields_ = [("x0", ctypes.c_double), ("y0", ctypes.c_double), ("rstar", ctypes.c_double), ("eps", ctypes.c_double)]


class VortexParamsC(ctypes.Structure):
    _fields_ = [("nvortices", ctypes.c_int32), ("pad_", ctypes.c_int32), ("vortices", Vortice * 5),
                ("mach", ctypes.c_double)]


_lib = None


def lib():
    """The loaded library (raises if the HIP extension is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libsbpgpu.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    c_int, c_double, c_int64, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_int64, ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    sig = {
        "sbp_status_string": ([c_int], ctypes.c_char_p),
        "sbp_set_gamma": ([c_double], c_int),
        "sbp_get_gamma": ([dp], c_int),
        "sbp_operator_info": ([c_int, ip, ip, ip, ip, ip], c_int),
        "sbp_operator_h": ([c_int, dp], c_int),
        "sbp_operator_coefficients": ([c_int, c_int, dp, dp, dp], c_int),
        "sbp_context_create": ([c_int, ctypes.POINTER(c_void_p)], c_int),
        "sbp_context_destroy": ([c_void_p], c_int),
        "sbp_context_set_stream": ([c_void_p, c_void_p], c_int),
        "sbp_context_synchronize": ([c_void_p], c_int),
        "sbp_diffxi": ([c_void_p, c_int, c_void_p, c_int64, c_void_p, c_int64, c_int, c_int], c_int),
        "sbp_diffeta": ([c_void_p, c_int, c_void_p, c_int64, c_void_p, c_int64, c_int, c_int], c_int),
        "sbp_dissxi": ([c_void_p, c_int, c_void_p, c_int64, c_void_p, c_int64, c_int, c_int], c_int),
        "sbp_disseta": ([c_void_p, c_int, c_void_p, c_int64, c_void_p, c_int64, c_int, c_int], c_int),
        "sbp_rhs": ([c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                     ctypes.POINTER(c_void_p)], c_int),
        "sbp_metrics": ([c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p], c_int),
        "sbp_system_create": ([c_void_p, c_int, ip, ip, ip, ip, ctypes.POINTER(dp), ctypes.POINTER(dp),
                               ctypes.POINTER(c_void_p)], c_int),
        "sbp_system_destroy": ([c_void_p], c_int),
        "sbp_system_set_bc": ([c_void_p, c_int, c_int, c_int, c_int, ctypes.POINTER(VortexParamsC)], c_int),
        "sbp_system_set_rhs_kind": ([c_void_p, c_int], c_int),
        "sbp_system_set_precision": ([c_void_p, c_int], c_int),
        "sbp_system_set_state": ([c_void_p, c_int, dp], c_int),
        "sbp_system_get_state": ([c_void_p, c_int, dp], c_int),
        "sbp_system_state_device": ([c_void_p, c_int, ctypes.POINTER(c_void_p), ctypes.POINTER(c_int64),
                                     ctypes.POINTER(c_int64)], c_int),
        "sbp_system_get_metrics": ([c_void_p, c_int, dp], c_int),
        "sbp_system_vortex": ([c_void_p, c_double, ctypes.POINTER(VortexParamsC)], c_int),
        "sbp_system_vortex_block": ([c_void_p, c_int, c_double, ctypes.POINTER(VortexParamsC)], c_int),
        "sbp_system_set_dt_scale": ([c_void_p, c_int, c_double], c_int),
        "sbp_system_rhs": ([c_void_p, c_double, dp], c_int),
        "sbp_system_advance": ([c_void_p, c_double, c_int64], c_int),
        "sbp_system_get_time": ([c_void_p, dp], c_int),
        "sbp_system_set_time": ([c_void_p, c_double], c_int),
        "sbp_system_max_dt": ([c_void_p, dp], c_int),
        "sbp_system_error_vortex": ([c_void_p, ctypes.POINTER(VortexParamsC), dp], c_int),
        "sbp_system_error_vortex_block": ([c_void_p, c_int, ctypes.POINTER(VortexParamsC), dp], c_int),
        "sbp_system_energy": ([c_void_p, dp], c_int),
        "sbp_system_check": ([c_void_p], c_int),
        "sbp_system_profile": ([c_void_p, c_double, c_int64, dp, dp], c_int),
        "sbp_system_tile_shape": ([c_void_p, ip, ip, ip], c_int),
        "sbp_system_set_tile_shape": ([c_void_p, c_int], c_int),
        "sbp_system_tile_index": ([c_void_p, ip], c_int),
        "sbp_system_set_schedule": ([c_void_p, c_int], c_int),
        "sbp_system_schedule": ([c_void_p, ip], c_int),
        "sbp_system_tile_order": ([c_void_p, ip], c_int),
        "sbp_system_set_tile_order": ([c_void_p, c_int], c_int),
        "sbp_system_set_bc_interpolate": ([c_void_p, c_int, c_int, c_int, c_int], c_int),
        "sbp_system_advance_adaptive": ([c_void_p, c_double, dp, c_double], c_int),
        "sbp_system_set_bc_multigrid": ([c_void_p, c_int, c_int, c_int, ip], c_int),
        "sbp_interpolate": ([c_void_p, c_int, c_int, c_void_p, c_int, c_void_p, c_int], c_int),
        "sbp_interp_info": ([c_int, c_int, ip], c_int),
        "sbp_interp_coefficients": ([c_int, c_int, dp, dp], c_int),
        "sbp_system_advance_timed": ([c_void_p, c_double, c_int64, dp], c_int),
        "sbp_rhs_no_sat": ([c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p], c_int),
        "sbp_sat_characteristics": ([c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                     ctypes.POINTER(c_void_p)], c_int),
        "sbp_sat_edge": ([c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p], c_int),
        "sbp_h2_err": ([c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p, dp], c_int),
        "sbp_h2linspace": ([c_double, c_double, c_int, dp], c_int),
        "sbp_h2_err_noalloc": ([c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p, dp], c_int),
        "sbp_expr_create": ([c_int, ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(c_void_p),
                             ctypes.c_char_p, c_int], c_int),
        "sbp_expr_destroy": ([c_void_p], c_int),
        "sbp_expr_eval_host": ([c_void_p, c_double, c_int64, dp, dp, dp], c_int),
        "sbp_system_set_bc_eval": ([c_void_p, c_int, c_int, c_void_p], c_int),
        "sbp_system_eval_block": ([c_void_p, c_int, c_double, c_void_p], c_int),
        "sbp_system_error_eval": ([c_void_p, c_void_p, dp], c_int),
        "sbp_system_integrate": ([c_void_p, c_int, dp, dp, dp, c_double, c_int64], c_int),
        "sbp_tile_shape_count": ([], c_int),
        "sbp_tile_shape_get": ([c_int, ip, ip, ip], c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res


EXPORTED = [
    "sbp_status_string", "sbp_set_gamma", "sbp_get_gamma", "sbp_operator_info", "sbp_operator_h",
    "sbp_operator_coefficients", "sbp_context_create", "sbp_context_destroy", "sbp_context_set_stream",
    "sbp_context_synchronize", "sbp_diffxi", "sbp_diffeta", "sbp_dissxi", "sbp_disseta", "sbp_rhs", "sbp_metrics",
    "sbp_system_create", "sbp_system_destroy", "sbp_system_set_bc", "sbp_system_set_rhs_kind",
    "sbp_system_set_precision", "sbp_system_set_state", "sbp_system_get_state", "sbp_system_state_device",
    "sbp_system_get_metrics", "sbp_system_vortex", "sbp_system_vortex_block", "sbp_system_set_dt_scale",
    "sbp_system_error_vortex_block", "sbp_system_rhs", "sbp_system_advance", "sbp_system_get_time",
    "sbp_system_set_time", "sbp_system_max_dt", "sbp_system_error_vortex", "sbp_system_energy", "sbp_system_check",
    "sbp_system_profile", "sbp_system_tile_shape", "sbp_system_set_tile_shape", "sbp_system_tile_index", "sbp_system_set_schedule", "sbp_system_schedule", "sbp_system_tile_order", "sbp_system_set_tile_order", "sbp_system_advance_timed",
    "sbp_system_set_bc_interpolate", "sbp_system_set_bc_multigrid", "sbp_interpolate", "sbp_interp_info",
    "sbp_system_advance_adaptive",
    "sbp_interp_coefficients",
    "sbp_tile_shape_count",
    "sbp_tile_shape_get",
    "sbp_rhs_no_sat", "sbp_sat_characteristics", "sbp_sat_edge", "sbp_h2_err", "sbp_h2linspace",
    "sbp_system_integrate", "sbp_h2_err_noalloc", "sbp_expr_create", "sbp_expr_destroy", "sbp_expr_eval_host",
    "sbp_system_set_bc_eval", "sbp_system_eval_block", "sbp_system_error_eval",
]


def status_string(status):
    return lib().sbp_status_string(status).decode()


def check(status, what):
    if status != OK:
        raise SbpError(status, what)


def wait_for_torch():
    """Wait for the work queued on torch's current stream. The library runs on
    its context's own stream, so torch tensors handed over by pointer must be
    finished (filled, copied) before a library call reads or writes them."""
    import torch

    torch.cuda.current_stream().synchronize()


_context = None


def context():
    """The process-wide context on the current HIP device (created on first use)."""
    global _context
    if _context is None:
        h = ctypes.c_void_p()
        check(lib().sbp_context_create(0, ctypes.byref(h)), "sbp_context_create")
        _context = h
    return _context


def vortex_params_c(vortices, mach):
    p = VortexParamsC()
    if len(vortices) > 5:
        raise ValueError("at most 5 vortices")
    p.nvortices = len(vortices)
    for i, v in enumerate(vortices):
        p.vortices[i].x0, p.vortices[i].y0, p.vortices[i].rstar, p.vortices[i].eps = v
    p.mach = mach
    return p
