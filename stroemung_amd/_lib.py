"""ctypes binding of libstroemung_hip.so (declarations in include/stroemung_hip.h).

The library is built in-tree (python -m stroemung_amd.build).  There is no
CPU fallback: if the library is missing, loading fails with an error.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libstroemung_hip.so")

STM_OK = 0
ERRORS = {
    -1: "NO_DEVICE", -2: "HIP", -3: "INVALID_GRID", -4: "STALE_BOUNDARY", -5: "NAN",
    -6: "INVALID_ARG", -7: "STATE",
}
SOLVER_LEX, SOLVER_RB = 0, 1


class StmError(RuntimeError):
    def __init__(self, code, message):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, message))
        self.code = code
        self.message = message


class NoDeviceError(StmError):
    pass


class InvalidGridError(StmError):
    """The reference's SimulationGridError::BoundaryTooThinError."""


class NanError(StmError):
    """A channel's pressure residual is not a number after a tick (the
    state is left as computed)."""


class Params(C.Structure):
    _fields_ = [
        ("delx", C.c_double), ("dely", C.c_double), ("delt", C.c_double),
        ("gamma", C.c_double), ("reynolds", C.c_double),
        ("sor_absolute_epsilon", C.c_double), ("max_iterations", C.c_uint32),
        ("omega", C.c_double), ("has_initial_norm_squared", C.c_int32),
        ("initial_norm_squared", C.c_double), ("iterations", C.c_uint32),
        ("time", C.c_double), ("tau", C.c_double), ("solver", C.c_int32),
        ("check_every", C.c_int32),
    ]


class State(C.Structure):
    _fields_ = [
        ("delt", C.c_double), ("time", C.c_double), ("initial_norm_squared", C.c_double),
        ("norm_squared", C.c_double), ("residual_squared", C.c_double),
        ("tolerance_squared", C.c_double), ("rhs_mean", C.c_double),
        ("pressure_range", C.c_double * 2),
        ("speed_range", C.c_double * 2), ("umax", C.c_double), ("vmax", C.c_double),
        ("iterations", C.c_uint32), ("sor_iterations", C.c_uint32),
        ("has_initial_norm_squared", C.c_int32), ("exhausted", C.c_int32),
    ]


class NetStats(C.Structure):
    _fields_ = [
        ("forward_ms", C.c_double), ("train_ms", C.c_double),
        ("flops_forward_per_sample", C.c_double),
        ("forward_calls", C.c_uint64), ("forward_samples", C.c_uint64),
        ("train_steps", C.c_uint64), ("train_samples", C.c_uint64),
        ("params", C.c_uint64), ("dataset_size", C.c_uint32), ("pad", C.c_uint32),
    ]


_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_NET = C.c_void_p
_U8 = C.POINTER(C.c_uint8)
_CTX = C.c_void_p

# name -> (restype, argtypes); the complete exported surface of the header.
SIGNATURES = {
    "stm_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "stm_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "stm_destroy": (None, [_CTX]),
    "stm_last_error": (C.c_char_p, [_CTX]),
    "stm_rebuild_boundary_list": (C.c_int, [C.c_uint32, C.c_uint32, _U8, C.POINTER(C.c_int8),
                                            C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), _D,
                                            C.c_char_p, C.c_size_t]),
    "stm_add_grid": (C.c_int, [_CTX, C.POINTER(Params), C.c_uint32, C.c_uint32, _U8, _D,
                               C.POINTER(C.c_uint32)]),
    "stm_set_cells": (C.c_int, [_CTX, C.c_uint32, _U8, _D]),
    "stm_set_params": (C.c_int, [_CTX, C.c_uint32, C.POINTER(Params)]),
    "stm_upload": (C.c_int, [_CTX, C.c_uint32] + [_D] * 6),
    "stm_download": (C.c_int, [_CTX, C.c_uint32] + [_D] * 6),
    "stm_get_state": (C.c_int, [_CTX, C.c_uint32, C.POINTER(State)]),
    "stm_construct": (C.c_int, [_CTX]),
    "stm_run_simulation_tick": (C.c_int, [_CTX, C.POINTER(C.c_uint32), _D]),
    "stm_run_ticks": (C.c_int, [_CTX, C.c_uint32]),
    "stm_set_boundary_u_and_v": (C.c_int, [_CTX]),
    "stm_calculate_f_and_g": (C.c_int, [_CTX]),
    "stm_calculate_rhs": (C.c_int, [_CTX]),
    "stm_solve_pressure": (C.c_int, [_CTX]),
    "stm_set_u_and_v": (C.c_int, [_CTX]),
    "stm_copy_pressure_to_boundaries": (C.c_int, [_CTX]),
    "stm_rb_sweeps": (C.c_int, [_CTX, C.c_uint32]),
    "stm_solver_variant": (C.c_int, [_CTX, C.c_int32] + [C.POINTER(C.c_int32)] * 4),
    "stm_autotune": (C.c_int, [_CTX, C.c_uint32, C.POINTER(C.c_int32), _D]),
    "stm_autotune_ticks": (C.c_int, [_CTX, C.POINTER(C.c_int32), _D]),
    "stm_synchronize": (C.c_int, [_CTX]),
    "stm_solver_bytes": (C.c_int, [_CTX, _D, _D]),
    "stm_stream_copy": (C.c_int, [_CTX, C.c_uint64, C.c_uint32, _D]),
    "stm_sync_benchmark": (C.c_int, [_CTX, C.c_uint32, _D]),
    "stm_profile": (C.c_int, [_CTX, C.c_uint32]),
    "stm_kernel_stats": (C.c_int, [_CTX, C.c_char_p, _D, C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_uint64)]),
    "stm_reset_stats": (C.c_int, [_CTX]),
    "stm_net_create": (C.c_int, [_CTX, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64,
                                 C.POINTER(C.c_void_p)]),
    "stm_net_destroy": (None, [_NET]),
    "stm_net_last_error": (C.c_char_p, [_NET]),
    "stm_net_param_count": (C.c_int, [_NET, C.POINTER(C.c_uint64)]),
    "stm_net_get_params": (C.c_int, [_NET, _F]),
    "stm_net_set_params": (C.c_int, [_NET, _F]),
    "stm_net_forward": (C.c_int, [_NET, C.c_uint32, _F, _F]),
    "stm_net_gradients": (C.c_int, [_NET, C.c_uint32, _F, _F, _F, _F]),
    "stm_net_dataset_reserve": (C.c_int, [_NET, C.c_uint32]),
    "stm_net_dataset_size": (C.c_int, [_NET, C.POINTER(C.c_uint32)]),
    "stm_net_dataset_get": (C.c_int, [_NET, C.c_uint32, _F, _F]),
    "stm_net_train": (C.c_int, [_NET, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64, _F]),
    "stm_net_stats": (C.c_int, [_NET, C.POINTER(NetStats)]),
    "stm_net_benchmark": (C.c_int, [_NET, C.c_uint32, C.c_uint32, _D, _D]),
    "stm_set_learned": (C.c_int, [_CTX, _NET, C.c_uint32, C.c_uint32, C.c_int32, C.c_uint32]),
    "stm_set_capture": (C.c_int, [_CTX, _NET, C.c_uint32, C.c_uint32, C.c_uint32]),
    "stm_set_graphs": (C.c_int, [_CTX, C.c_int32]),
    "stm_render": (C.c_int, [_CTX, C.c_uint32, C.c_int32, _U8]),
    "stm_render_static": (C.c_int, [_CTX, C.c_uint32, C.c_int32, _U8]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libstroemung_hip.so is not built (%s); run "
                              "`python -m stroemung_amd.build`" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc, ctx=None, net=None):
    if rc == STM_OK:
        return
    if net is not None:
        msg = lib().stm_net_last_error(net).decode(errors="replace")
    else:
        msg = lib().stm_last_error(ctx).decode(errors="replace")
    if rc == -1:
        raise NoDeviceError(rc, msg)
    if rc == -3:
        raise InvalidGridError(rc, msg)
    if rc == -5:
        raise NanError(rc, msg)
    raise StmError(rc, msg)
