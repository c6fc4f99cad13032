"""ctypes binding of libfestung_hip.so (the C-ABI declared in
include/festung_hip.h).

The library is loaded from this package directory only; there is no
fallback: if it is missing the import fails with instructions to build it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libfestung_hip.so')

OK, ERR_ARG, ERR_HIP, ERR_ALLOC, ERR_STATE, ERR_UNSUPPORTED, ERR_INTERNAL = range(7)
FLAG_LIMITER = 1
FLAG_GRAPH = 2
FLAG_ADJ_SERIAL = 4
FLAG_ADJ_FUSED = 8
FLAG_ADJ_SPLIT = 16
FLAG_ADJ_DEFER = 32
FLAG_LIM_HIERARCH = 64
FLAG_LIM_STRICT = 128
FLAG_ADJ_BINOMIAL = 256
OP_G, OP_G1, OP_G2, OP_R, OP_M, OP_MINV = range(6)
LIMITER_TYPES = {'linear': 0, 'hierarch_vert': FLAG_LIM_HIERARCH, 'strict': FLAG_LIM_STRICT}


class FestungError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, code, msg):
        super().__init__('festung error %d: %s' % (code, msg))
        self.code = code


_lib = None

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int32)
vp = C.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    'festung_last_error': (C.c_char_p, []),
    'festung_version': (C.c_char_p, []),
    'festung_domain_square': (C.c_int, [C.c_double, C.c_double, C.c_double, C.POINTER(vp)]),
    'festung_grid_create': (C.c_int, [dp, C.c_int32, ip, C.c_int32, C.POINTER(vp)]),
    'festung_grid_set_boundary': (C.c_int, [vp, ip, C.POINTER(C.c_uint8)]),
    'festung_grid_concat': (C.c_int, [C.POINTER(vp), C.c_int32, C.POINTER(vp)]),
    'festung_grid_sizes': (C.c_int, [vp, ip, ip, ip]),
    'festung_grid_field_size': (C.c_int, [vp, C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    'festung_grid_get_i32': (C.c_int, [vp, C.c_char_p, ip]),
    'festung_grid_get_f64': (C.c_int, [vp, C.c_char_p, dp]),
    'festung_grid_map_ref_to_phy': (C.c_int, [vp, C.c_int32, dp, dp, dp, dp]),
    'festung_grid_destroy': (None, [vp]),
    'festung_ref_size': (C.c_int, [C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_int64)]),
    'festung_phi_at_points': (C.c_int, [C.c_int, C.c_int32, dp, dp, dp]),
    'festung_ref_get': (C.c_int, [C.c_int, C.c_char_p, C.c_int, dp]),
    'festung_problem_create': (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]),
    'festung_problem_destroy': (None, [vp]),
    'festung_problem_sizes': (C.c_int, [vp, ip, ip, ip, ip, ip]),
    'festung_problem_stream': (C.c_int, [vp, C.POINTER(vp)]),
    'festung_set_velocity_disc': (C.c_int, [vp, dp, dp, dp]),
    'festung_set_velocity_points': (C.c_int, [vp, dp, dp, dp, dp]),
    'festung_set_source_points': (C.c_int, [vp, dp, dp, dp]),
    'festung_set_vertex_dirichlet': (C.c_int, [vp, dp]),
    'festung_set_velocity_points_dev': (C.c_int, [vp, vp, vp, vp, vp]),
    'festung_set_source_points_dev': (C.c_int, [vp, vp, vp, vp]),
    'festung_set_vertex_dirichlet_dev': (C.c_int, [vp, vp]),
    'festung_project': (C.c_int, [vp, C.c_int, dp, dp, C.c_int]),
    'festung_set_state': (C.c_int, [vp, dp]),
    'festung_get_state': (C.c_int, [vp, dp]),
    'festung_set_state_dev': (C.c_int, [vp, vp]),
    'festung_get_state_dev': (C.c_int, [vp, vp]),
    'festung_begin_step': (C.c_int, [vp]),
    'festung_stage': (C.c_int, [vp, C.c_int, C.c_double]),
    'festung_step': (C.c_int, [vp, C.c_int, C.c_double]),
    'festung_limit': (C.c_int, [vp]),
    'festung_limit_disc_linear': (C.c_int, [vp]),
    'festung_rhs': (C.c_int, [vp, dp]),
    'festung_limiter_info': (C.c_int, [vp, dp, dp, dp]),
    'festung_l2_error': (C.c_int, [vp, C.c_int, dp, dp]),
    'festung_set_tau_field': (C.c_int, [vp, dp]),
    'festung_set_elem_block': (C.c_int, [vp, C.c_int32]),
    'festung_get_elem_block': (C.c_int, [vp, C.POINTER(C.c_int32)]),
    'festung_autotune': (C.c_int, [vp, C.c_double, C.c_int32, C.POINTER(C.c_int32), dp]),
    'festung_tune_save': (C.c_int, [C.c_char_p]),
    'festung_tune_load': (C.c_int, [C.c_char_p]),
    'festung_tune_clear': (C.c_int, []),
    'festung_tune_count': (C.c_int, [C.POINTER(C.c_int32)]),
    'festung_batch_create': (C.c_int, [C.POINTER(vp), C.c_int32, C.POINTER(vp)]),
    'festung_batch_step': (C.c_int, [vp, C.c_int32, dp]),
    'festung_batch_sync': (C.c_int, [vp]),
    'festung_batch_destroy': (None, [vp]),
    'festung_adjoint': (C.c_int, [vp, C.c_int32, C.c_double, dp, C.c_int32, dp, dp, dp, dp, dp, dp]),
    'festung_binomial_schedule': (C.c_int, [C.c_int32, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    'festung_adj_rhs': (C.c_int, [vp, dp, dp]),
    'festung_adjoint_dev': (C.c_int, [vp, C.c_int32, C.c_double, vp, C.c_int32, dp, vp, vp, vp, vp, dp]),
    'festung_apply_operator': (C.c_int, [vp, C.c_int, dp, dp]),
    'festung_apply_operator_dev': (C.c_int, [vp, C.c_int, vp, vp]),
    'festung_apply_operator_dev_noalloc': (C.c_int, [vp, C.c_int, vp, vp]),
    'festung_assemble_vectors': (C.c_int, [vp, dp, dp, dp, dp, dp, dp]),
    'festung_vertex_based_limiter': (C.c_int, [vp, dp, dp, C.POINTER(C.c_uint8), dp, dp, dp, dp]),
    'festung_apply_slope_limiter_disc': (C.c_int, [vp, dp, C.POINTER(C.c_uint8), dp, dp, dp, dp]),
    'festung_runge_kutta_explicit': (C.c_int, [C.c_int, C.c_double, C.c_double, dp, dp, ip]),
    'festung_profile': (C.c_int, [vp, C.c_int]),
    'festung_profile_read': (C.c_int, [vp, dp, C.POINTER(C.c_int64)]),
    'festung_sync': (C.c_int, [vp]),
}


def lib():
    """The loaded library (loads it on first use; raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError('%s is not built; run `python -m festung_amd.build` '
                              '(or __graft_entry__.build())' % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != OK:
        msg = lib().festung_last_error().decode(errors='replace')
        raise FestungError(status, msg)


def as_f64(a, shape=None):
    """C-contiguous float64 view (copy if needed) of a column-major (Matlab)
    array given as its transpose, i.e. numpy array laid out [..., K]."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != tuple(shape):
        raise ValueError('expected shape %s, got %s' % (shape, a.shape))
    return a


def ptr(a):
    if a is None:
        return None
    if a.dtype == np.float64:
        return a.ctypes.data_as(dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(ip)
    if a.dtype == np.uint8:
        return a.ctypes.data_as(C.POINTER(C.c_uint8))
    raise TypeError(a.dtype)
