"""ctypes binding of libmuscade_hip.so (declared in include/muscade_hip.h).

The library is the only compute path: if it cannot be loaded, importing this
module raises. There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libmuscade_hip.so")

MH_OK, MH_ERR_ARG, MH_ERR_HIP, MH_ERR_NAN, MH_ERR_ORIENT, MH_ERR_SIZE, MH_ERR_SOLVE = 0, 1, 2, 3, 4, 5, 6
MH_KIND_BEAM, MH_KIND_BEAM_U, MH_KIND_HOLD, MH_KIND_DOFLOAD, MH_KIND_ZERO, MH_KIND_BAR, MH_KIND_BAR_U = 1, 2, 3, 4, 5, 6, 7
MH_KIND_HOST = 8

# every function of include/muscade_hip.h, with its ctypes signature
_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.POINTER(ctypes.c_int64)
_I = ctypes.c_int
_L = ctypes.c_int64


class mh_err(ctypes.Structure):
    _fields_ = [("code", ctypes.c_int32), ("igroup", ctypes.c_int32), ("iele", ctypes.c_int64)]


SIGNATURES = {
    "mh_ctx_create": (_I, [_I, ctypes.POINTER(_P)]),
    "mh_ctx_destroy": (_I, [_P]),
    "mh_last_error": (ctypes.c_char_p, [_P]),
    "mh_strerror": (ctypes.c_char_p, [_I]),
    "mh_model_begin": (_I, [_P, _L, _L, _L, _L, _D]),
    "mh_group_add": (_I, [_P, _I, _L, _I, _I64, _D, _I, _I64, _I64, _D, _D, ctypes.POINTER(_I)]),
    "mh_group_set_adofs": (_I, [_P, _I, _I64, _I]),
    "mh_beam_dRdA": (_I, [_P, _I, _D]),
    "mh_adjoint_solve": (_I, [_P, _D, ctypes.c_double]),
    "mh_adjoint_solve_sparse": (_I, [_P, _L, _I64, _D, ctypes.c_double]),
    "mh_prepare_sweepx": (_I, [_P, _I, _I64]),
    "mh_get_pattern": (_I, [_P, _I64, _I64]),
    "mh_get_asm": (_I, [_P, _I, _I64, _I64]),
    "mh_set_X": (_I, [_P, _I, _D]),
    "mh_set_U": (_I, [_P, _D]),
    "mh_set_A": (_I, [_P, _D]),
    "mh_set_point_values": (_I, [_P, _I, _D]),
    "mh_set_time": (_I, [_P, ctypes.c_double]),
    "mh_set_lambda": (_I, [_P, _D]),
    "mh_vjp_lagrangian": (_I, [_P, _I, ctypes.POINTER(mh_err)]),
    "mh_get_vjp": (_I, [_P, _D, _D, _D]),
    "mh_get_tuning": (_I, [_P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float)]),
    "mh_get_beam_tuning": (_I, [_P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float)]),
    "mh_set_scatter_variant": (_I, [_P, _I]),
    "mh_set_group_scales": (_I, [_P, _I, _D, _D]),
    "mh_prepare_direct": (_I, [_P, _I, _I, _I, _I64]),
    "mh_get_direct_pattern": (_I, [_P, _I, _I64, _I64]),
    "mh_assemble_direct": (_I, [_P, _I, ctypes.c_double, ctypes.POINTER(mh_err)]),
    "mh_get_direct": (_I, [_P, _D, _D]),
    "mh_assemble_sweepx": (_I, [_P, _I, _I, _D, ctypes.POINTER(mh_err)]),
    "mh_assemble_sweepx_graph": (_I, [_P, _I, _I, _D, ctypes.POINTER(mh_err)]),
    "mh_graph_captures": (_I, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "mh_get_rhs": (_I, [_P, _D]),
    "mh_get_nzval": (_I, [_P, _D]),
    "mh_device_outputs": (_I, [_P, ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "mh_get_element_blocks": (_I, [_P, _I, _D, _D]),
    "mh_chain_prepare": (_I, [_P, _I, _L, _I64, _L, _I64, _L, _I64]),
    "mh_chain_solve": (_I, [_P]),
    "mh_chain_timings": (_I, [_P, _D]),
    "mh_get_dx": (_I, [_P, _D]),
    "mh_newton_prepare": (_I, [_P, _D]),
    "mh_newton_update": (_I, [_P, _I, _I, _D, _D]),
    "mh_newton_iteration_graph": (_I, [_P, _I, _I, _D, _D, ctypes.POINTER(mh_err)]),
    "mh_get_X": (_I, [_P, _I, _D]),
    "mh_beam_results": (_I, [_P, _I, _I, _I, _D, _D]),
    "mh_set_host_blocks": (_I, [_P, _I, _D, _D]),
    "mh_set_host_dr": (_I, [_P, _I, _D]),
    "mh_big_prepare": (_I, [_P, _L, _L, _I64, _I64, _I64, _D, _L, _I64]),
    "mh_big_get_pattern": (_I, [_P, _I64, _I64]),
    "mh_big_assemble": (_I, [_P, _D, _D]),
    "mh_big_solve_prepare": (_I, [_P, _L, _L, _L, _I64, _I64]),
    "mh_big_solve": (_I, [_P, _D, _D]),
    "mh_beam_gauge_derivs": (_I, [_P, _I, _L, _I64, _D]),
    "mh_beam_gauge_derivs_noalloc": (_I, [_P, _I, _L, _I64, _D]),
    "mh_last_timings": (_I, [_P, _D]),
    "mh_calib_copy": (_I, [_P, _L, _D]),
    "mh_sync": (_I, [_P]),
}


def _load():
    if not os.path.exists(LIBPATH):
        raise RuntimeError(
            f"{LIBPATH} is missing: build it with __graft_entry__.build() (hipcc, gfx950). "
            "There is no CPU fallback for the element-assembly path.")
    lib = ctypes.CDLL(LIBPATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def dptr(a):
    return a.ctypes.data_as(_D) if a is not None else None


def i64ptr(a):
    return a.ctypes.data_as(_I64) if a is not None else None


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def as_i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)
