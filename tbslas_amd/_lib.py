"""ctypes binding of libtbx.so (the C ABI declared in include/tbx.h).

The library is the only compute path: if it is not built this module raises
instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtbx.so")

TBX_OK = 0
BC_FREESPACE = 0
BC_PERIODIC = 1
FP64 = 0x0
EXACT = 0x1
FP32 = 0x2
FP32_HEAD64 = 0x3
WRITE_BACK_WRAP = 0x10
MAX_Q = 16
MAX_DEPTH = 21
OPT_STEP_SHAPE = 1
OPT_TRACER_MFMA = 2
OPT_GRID_LDS = 3
OPT_NODE_VELOCITY = 4
VEL_TREE = 0
VEL_EXTRAP = 1
VEL_SET = 2

MODES = {"fp64": FP64, "exact": EXACT, "fp32": FP32, "fp32h": FP32_HEAD64}

TFAC_FN = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)


class TbxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"tbx error {code}: {msg}")
        self.code = code


class _TreeDesc(C.Structure):
    _fields_ = [
        ("num_leaves", C.c_int64),
        ("anchor", C.c_void_p),
        ("depth", C.c_void_p),
        ("q", C.c_int32),
        ("dof", C.c_int32),
        ("coeff", C.c_void_p),
    ]


class _GridDesc(C.Structure):
    _fields_ = [
        ("num_leaves", C.c_int64),
        ("anchor", C.c_void_p),
        ("depth", C.c_void_p),
        ("n", C.c_int32),
        ("dof", C.c_int32),
        ("samples", C.c_void_p),
    ]


class _Velocity(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("trees", C.c_void_p * 4),
        ("times", C.c_double * 4),
        ("g", TFAC_FN),
        ("gctx", C.c_void_p),
    ]


_lib = None

_P = C.c_void_p
_I64 = C.c_int64
_I32 = C.c_int
_D = C.c_double
_U32 = C.c_uint32

_SIGS = {
    "tbx_tree_create": (_I32, [C.POINTER(_TreeDesc), C.POINTER(_P)]),
    "tbx_tree_destroy": (_I32, [_P]),
    "tbx_tree_info": (_I32, [_P, _P, _P, _P, _P]),
    "tbx_tree_set_coeffs": (_I32, [_P, _P, _I32, _P]),
    "tbx_tree_get_coeffs": (_I32, [_P, _P, _I32, _P]),
    "tbx_locate": (_I32, [_P, _P, _I64, _I32, _P, _P]),
    "tbx_eval": (_I32, [_P, _P, _I64, _I32, _U32, _P, _P]),
    "tbx_traj_rk2": (_I32, [_P, TFAC_FN, _P, _P, _I64, _D, _D, _I32, _I32, _U32, _P, _P]),
    "tbx_semilag_rk2": (_I32, [_P, TFAC_FN, _P, _P, _P, _I64, _I32, _D, _I32, _I32, _U32,
                               _P, _P]),
    "tbx_collect_grid_points": (_I32, [_P, _P, _P]),
    "tbx_set_grid_values": (_I32, [_P, _P, _P]),
    "tbx_tree_grid_values": (_I32, [_P, _P, _P]),
    "tbx_tree_grid_values_f32": (_I32, [_P, _P, _P]),
    "tbx_semilag_insitu": (_I32, [_P, TFAC_FN, _P, _P, _I32, _D, _I32, _I32, _U32, _P, _P]),
    "tbx_semilag_values": (_I32, [_P, TFAC_FN, _P, _P, _I32, _D, _I32, _I32, _U32, _P, _P, _P]),
    "tbx_semilag_insitu_noalloc": (_I32, [_P, TFAC_FN, _P, _P, _I32, _D, _I32, _I32, _U32, _P,
                                          C.POINTER(C.c_size_t), _P]),
    "tbx_velocity_eval": (_I32, [C.POINTER(_Velocity), _P, _I64, _D, _I32, _U32, _P, _P]),
    "tbx_traj_rk2_v": (_I32, [C.POINTER(_Velocity), _P, _I64, _D, _D, _I32, _I32, _U32, _P,
                              _P]),
    "tbx_semilag_values_v": (_I32, [C.POINTER(_Velocity), _P, _I32, _D, _I32, _I32, _U32, _P,
                                    _P, _P]),
    "tbx_semilag_insitu_v": (_I32, [C.POINTER(_Velocity), _P, _I32, _D, _I32, _I32, _U32, _P,
                                    _P]),
    "tbx_semilag_insitu_vjp": (_I32, [_P, TFAC_FN, _P, _P, _I32, _D, _I32, _I32, _U32, _P, _P,
                                      _P, _P, _P]),
    "tbx_adjoint_run": (_I32, [_P, TFAC_FN, _P, _P, _I32, _D, _I32, _U32, _P, _I32, _P, _P,
                               C.POINTER(C.c_double), _P]),
    "tbx_grid_create": (_I32, [C.POINTER(_GridDesc), C.POINTER(_P)]),
    "tbx_grid_destroy": (_I32, [_P]),
    "tbx_grid_set_samples": (_I32, [_P, _P, _I32, _P]),
    "tbx_grid_get_samples": (_I32, [_P, _P, _I32, _P]),
    "tbx_grid_eval": (_I32, [_P, _P, _I64, _I32, _U32, _P, _P]),
    "tbx_grid_semilag_insitu": (_I32, [_P, TFAC_FN, _P, _P, _I32, _D, _I32, _I32, _U32, _P, _P]),
    "tbx_tree_first_kind_points": (_I32, [_P, _P, _P]),
    "tbx_tree_fit": (_I32, [_P, _P, _P]),
    "tbx_error_points": (_I32, [_P, _P, _P]),
    "tbx_check_cheb_output": (_I32, [_P, _P, _U32, _P, _P]),
    "tbx_tree_write_vtu": (_I32, [_P, C.c_char_p, _I32, _P]),
    "tbx_tree_adapt": (_I32, [_P, _D, _I32, _I32, _I32, _P]),
    "tbx_tree_get_leaves": (_I32, [_P, _P, _P]),
    "tbx_batch_create": (_I32, [_I32, _P, _P, _P, _I32, C.POINTER(_P)]),
    "tbx_batch_destroy": (_I32, [_P]),
    "tbx_batch_set_config": (_I32, [_P, _I32]),
    "tbx_batch_get_config": (_I32, [_P, C.POINTER(_I32)]),
    "tbx_batch_node_velocity": (_I32, [_P, C.POINTER(_I32)]),
    "tbx_batch_tune": (_I32, [_P, _P, _P]),
    "tbx_batch_step": (_I32, [_P, _I32, _P]),
    "tbx_batch_attach_step": (_I32, [_P, _P, _P, _D, _I32, _U32]),
    "tbx_batch_get_coeffs": (_I32, [_P, _I32, _P, _P]),
    "tbx_new_nodes": (_I32, [_I32, _P]),
    "tbx_pt2coeff_matrix": (_I32, [_I32, _P]),
    "tbx_refit_gram_inverse": (_I32, [_I32, _P]),
    "tbx_uniform_leaves": (_I32, [_I32, _P, _P]),
    "tbx_first_kind_points": (_I32, [_I64, _P, _P, _I32, _P]),
    "tbx_fit_leaves": (_I32, [_I64, _I32, _I32, _P, _P]),
    "tbx_tail_norm": (_I32, [_I64, _I32, _I32, _P, _P]),
    "tbx_balance21": (_I32, [C.POINTER(_I64), _I64, _P, _P, _I32]),
    "tbx_set_option": (_I32, [_I32, _I32]),
    "tbx_get_option": (_I32, [_I32, C.POINTER(_I32)]),
    "tbx_step_shape": (_I32, [C.c_uint32, _I32, C.POINTER(_I32)]),
    "tbx_contraction_only": (_I32, [_P, _I64, _I32, C.c_uint32, _P, _P, _P]),
    "tbx_last_error": (C.c_char_p, []),
    "tbx_version": (C.c_char_p, []),
}


def lib() -> C.CDLL:
    """Load libtbx.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`"
            )
        # torch loads its own copy of the HIP/HSA runtime by path.  Loading it
        # first lets libtbx's libamdhip64.so.7 dependency bind to that copy;
        # the other order puts two HSA runtimes in the process and the second
        # one finds no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != TBX_OK:
        raise TbxError(rc, lib().tbx_last_error().decode())


def tfac(g):
    """Wrap a Python callable g(t) -> float as the C time-factor callback."""
    if g is None:
        return TFAC_FN()
    return TFAC_FN(lambda t, ctx: float(g(t)))


def set_option(option: int, value: int) -> None:
    check(lib().tbx_set_option(option, value))


def get_option(option: int) -> int:
    v = C.c_int()
    check(lib().tbx_get_option(option, C.byref(v)))
    return v.value


def step_shape(mode: str, q: int) -> int:
    """The launch shape a step in `mode` at degree q runs with now."""
    v = C.c_int()
    check(lib().tbx_step_shape(MODES[mode], q, C.byref(v)))
    return v.value
