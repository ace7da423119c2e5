"""ctypes binding of libqrochet_hip.so (declarations in include/qrochet_hip.h).

Loading fails loudly: there is no CPU fallback behind this package.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("QROCHET_HIP_LIB", os.path.join(os.path.dirname(_HERE), "libqrochet_hip.so"))

QH_OK = 0
QH_ERR_ARGUMENT = -1
QH_ERR_MISSING_SCHMIDT = -2
QH_ERR_DIMENSION = -3
QH_ERR_ASSERTION = -4
QH_ERR_OOM = -5
QH_ERR_HIP = -6
QH_ERR_NOT_CONVERGED = -7
QH_ERR_ADMISSION = -8
QH_ERR_NUMERIC = -9

QH_CANONICAL = 1
QH_RENORMALIZE = 2
QH_PREC_FP64 = 0
QH_PREC_FP32 = 16
QH_PREC_BF16X3 = 32
QH_LEFT = 0
QH_RIGHT = 1
QH_METHOD_QR = 0
QH_METHOD_SVD = 1


class QHObservable(C.Structure):
    _fields_ = [("nsites", C.c_int), ("site", C.c_int), ("op", C.POINTER(C.c_double))]


_ctx_p = C.c_void_p
_mps_p = C.c_void_p
_mpo_p = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

# name -> (restype, argtypes)
SIGNATURES = {
    "qh_ctx_create": (C.c_int, [C.c_int, C.POINTER(_ctx_p)]),
    "qh_ctx_destroy": (C.c_int, [_ctx_p]),
    "qh_ctx_synchronize": (C.c_int, [_ctx_p]),
    "qh_last_error": (C.c_char_p, [_ctx_p]),
    "qh_ctx_set_timing": (C.c_int, [_ctx_p, C.c_int]),
    "qh_ctx_get_timing": (C.c_int, [_ctx_p, C.c_char_p, _dp, C.POINTER(C.c_int64), _dp, _dp]),
    "qh_ctx_reset_timing": (C.c_int, [_ctx_p]),
    "qh_ctx_set_jacobi": (C.c_int, [_ctx_p, C.c_int, C.c_int]),
    "qh_ctx_set_qr_passes": (C.c_int, [_ctx_p, C.c_int]),
    "qh_ctx_mem_info": (C.c_int, [_ctx_p, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "qh_ctx_autotune": (C.c_int, [_ctx_p, C.c_int, C.c_int, _dp, _ip]),
    "qh_ctx_autotune_reset": (C.c_int, [_ctx_p]),
    "qh_ctx_set_graphs": (C.c_int, [_ctx_p, C.c_int]),
    "qh_ctx_set_pivoting": (C.c_int, [_ctx_p, C.c_int]),
    "qh_mps_create": (C.c_int, [_ctx_p, C.c_int, C.c_int, _ip, C.POINTER(_dp), C.POINTER(_dp), C.POINTER(_mps_p)]),
    "qh_pmps_create": (C.c_int, [_ctx_p, C.c_int, C.c_int, _ip, C.POINTER(_dp), C.POINTER(_mps_p)]),
    "qh_mps_rand": (C.c_int, [_ctx_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.POINTER(_mps_p)]),
    "qh_mps_is_periodic": (C.c_int, [_mps_p, _ip, _ip]),
    "qh_mps_clone": (C.c_int, [_ctx_p, _mps_p, C.POINTER(_mps_p)]),
    "qh_mps_free": (C.c_int, [_mps_p]),
    "qh_mps_nsites": (C.c_int, [_mps_p, _ip, _ip]),
    "qh_mps_bond_dims": (C.c_int, [_mps_p, _ip, _ip]),
    "qh_mps_get_site": (C.c_int, [_mps_p, C.c_int, _dp]),
    "qh_mps_get_lambda": (C.c_int, [_mps_p, C.c_int, _dp]),
    "qh_mps_footprint": (C.c_int, [_mps_p, C.POINTER(C.c_size_t)]),
    "qh_evolve1": (C.c_int, [_ctx_p, _mps_p, C.c_int, _dp]),
    "qh_evolve2": (C.c_int, [_ctx_p, _mps_p, C.c_int, _dp, C.c_double, C.c_int, C.c_uint, _ip, _dp]),
    "qh_evolve2_batch": (
        C.c_int,
        [_ctx_p, C.c_int, C.POINTER(_mps_p), _ip, _dp, _dp, _ip, C.c_uint, _ip, _dp],
    ),
    "qh_truncate": (C.c_int, [_ctx_p, _mps_p, C.c_int, C.c_double, C.c_int]),
    "qh_contract_between": (C.c_int, [_ctx_p, _mps_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "qh_canonize_site": (C.c_int, [_ctx_p, _mps_p, C.c_int, C.c_int, C.c_int]),
    "qh_canonize": (C.c_int, [_ctx_p, _mps_p]),
    "qh_mixed_canonize": (C.c_int, [_ctx_p, _mps_p, C.c_int]),
    "qh_normalize": (C.c_int, [_ctx_p, _mps_p, C.c_int]),
    "qh_is_left_canonical": (C.c_int, [_ctx_p, _mps_p, C.c_int, C.c_double, _ip, _dp]),
    "qh_is_right_canonical": (C.c_int, [_ctx_p, _mps_p, C.c_int, C.c_double, _ip, _dp]),
    "qh_expect": (C.c_int, [_ctx_p, _mps_p, C.c_int, C.POINTER(QHObservable), _dp]),
    "qh_expect_many": (C.c_int, [_ctx_p, _mps_p, C.c_int, _ip, C.POINTER(QHObservable), _dp]),
    "qh_overlap": (C.c_int, [_ctx_p, _mps_p, _mps_p, _dp]),
    "qh_norm": (C.c_int, [_ctx_p, _mps_p, _dp]),
    "qh_overlap_noalloc": (C.c_int, [_ctx_p, _mps_p, _mps_p, _dp]),
    "qh_mpo_create": (C.c_int, [_ctx_p, C.c_int, C.c_int, _ip, C.POINTER(_dp), C.POINTER(_mpo_p)]),
    "qh_pmpo_create": (C.c_int, [_ctx_p, C.c_int, C.c_int, _ip, C.POINTER(_dp), C.POINTER(_mpo_p)]),
    "qh_mpo_free": (C.c_int, [_mpo_p]),
    "qh_mpo_apply": (C.c_int, [_ctx_p, _mpo_p, _mps_p, C.POINTER(_mps_p)]),
    "qh_mpo_fit": (C.c_int, [_ctx_p, _mpo_p, _mps_p, _mps_p, C.c_int, C.c_uint, _dp]),
}

_lib = None


def load():
    """Load the shared library (raises OSError if it is missing or unloadable)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"libqrochet_hip.so not found at {LIB_PATH}; build it with __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
