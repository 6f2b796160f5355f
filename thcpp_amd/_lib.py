"""ctypes binding of libthcpp_amd.so (the C ABI declared in include/thcpp_amd.h).

The library must have been built (python -m thcpp_amd.build); loading fails
loudly otherwise -- there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libthcpp_amd.so")

THC_OK = 0
THC_ERR_ARG = 1
THC_ERR_HIP = 2
THC_ERR_EMPTY_CLUSTER = 3
THC_ERR_NOT_POSDEF = 4
THC_ERR_NO_MEMORY = 5
THC_ERR_UNSUPPORTED = 6
THC_WARN_DUPLICATE_INDICES = 100
THC_WARN_LUV_NOT_POSDEF = 101
THC_WARN_RANK_DEFICIENT = 102

PREC_FP64, PREC_FP32, PREC_MIXED = 0, 1, 3
GEMM_AUTO, GEMM_TINY, GEMM_SMALL, GEMM_LARGE, GEMM_DTL, GEMM_DEEPK = 0, 1, 2, 3, 4, 5

_vp = ctypes.c_void_p
_i = ctypes.c_int
_l = ctypes.c_long
_u = ctypes.c_uint
_d = ctypes.c_double


class KmeansStats(ctypes.Structure):
    _fields_ = [("iterations", _i), ("converged", _i), ("duplicates", _i), ("interp_zeroed", _i),
                ("last_diff", _d), ("diffs", ctypes.POINTER(_d)), ("pcts", ctypes.POINTER(_d)),
                ("elapsed_ms", _d)]


class IsdfOpts(ctypes.Structure):
    _fields_ = [("precision", _i), ("refine_steps", _i), ("want_luv", _i), ("check_rank", _i)]


class IsdfStats(ctypes.Structure):
    _fields_ = [("cct_chol_info", _i), ("luv_chol_info", _i), ("cct_rank", _i), ("cct_smax", _d),
                ("stage_ms", _d * 12)]


class System(ctypes.Structure):
    _fields_ = [("grid", _vp), ("rho", _vp), ("aoR", _vp), ("coulG", _vp), ("n", _l), ("m", _i),
                ("thc_cfac", _i), ("max_it", _i), ("threshold", _d), ("seed", _u), ("fft_grid", _i * 3),
                ("interp_idx", _vp), ("v", _vp), ("luv", _vp), ("status", _i), ("kstats", KmeansStats),
                ("istats", IsdfStats)]


# name -> (restype, argtypes); every entry point of include/thcpp_amd.h
SIGNATURES = {
    "thc_ctx_create": (_i, [_i, _vp, ctypes.POINTER(_vp)]),
    "thc_ctx_destroy": (_i, [_vp]),
    "thc_ctx_set_graphs": (_i, [_vp, _i]),
    "thc_status_string": (ctypes.c_char_p, [_i]),
    "thc_last_error": (ctypes.c_char_p, []),
    "thc_version": (ctypes.c_char_p, []),
    "thc_kmeans_initial_indices": (_i, [_l, _i, _u, _vp]),
    "thc_classify_grid_points": (_i, [_vp, _vp, _l, _vp, _i, _vp]),
    "thc_update_centroids": (_i, [_vp, _vp, _vp, _l, _vp, _i, _vp]),
    "thc_map_to_grid": (_i, [_vp, _vp, _l, _vp, _i, _vp]),
    "thc_kmeans": (_i, [_vp, _vp, _vp, _l, _i, _i, _d, _u, _vp, _vp, _vp, _vp, ctypes.POINTER(KmeansStats)]),
    "thc_czt": (_i, [_vp, _vp, _l, _i, _vp, _i, _vp]),
    "thc_cct": (_i, [_vp, _vp, _l, _vp, _i, _vp]),
    "thc_cholesky": (_i, [_vp, _vp, _i, ctypes.POINTER(_i)]),
    "thc_cct_rank": (_i, [_vp, _vp, _i, ctypes.POINTER(_i), ctypes.POINTER(_d)]),
    "thc_cholesky_noalloc": (_i, [_vp, _vp, _i, ctypes.POINTER(_i)]),
    "thc_device_sqrt": (_i, [_vp, _vp, _l, _vp]),
    "thc_gemm_tuning": (_i, [_vp, _i]),
    "thc_gemm_time": (_i, [_vp, _l, _l, _l, _i, _i, _i, ctypes.POINTER(_d)]),
    "thc_least_squares": (_i, [_vp, _vp, _vp, _l, _i, _i, _vp, ctypes.POINTER(_i)]),
    "thc_fft_vectors": (_i, [_vp, _vp, _i, ctypes.POINTER(_i), _vp, _vp]),
    "thc_construct_muv": (_i, [_vp, _vp, _i, _l, _vp]),
    "thc_isdf": (_i, [_vp, _vp, _l, _i, _vp, _i, ctypes.POINTER(_i), _vp, ctypes.POINTER(IsdfOpts), _vp, _vp,
                      _vp, ctypes.POINTER(IsdfStats)]),
    "thc_factorize": (_i, [_vp, _vp, _vp, _vp, _l, _i, _i, _i, _d, _u, ctypes.POINTER(_i), _vp,
                           ctypes.POINTER(IsdfOpts), _vp, _vp, _vp, ctypes.POINTER(KmeansStats),
                           ctypes.POINTER(IsdfStats)]),
    "thc_czt_complex": (_i, [_vp, _vp, _l, _i, _vp, _i, _vp]),
    "thc_isdf_complex": (_i, [_vp, _vp, _l, _i, _vp, _i, ctypes.POINTER(_i), _vp, ctypes.POINTER(IsdfOpts), _vp, _vp,
                              _vp, ctypes.POINTER(IsdfStats)]),
    "thc_qrcp_points": (_i, [_vp, _vp, _vp, _l, _i, _i, _i, _vp, _vp]),
    "thc_qrcp_points_complex": (_i, [_vp, _vp, _vp, _l, _i, _i, _i, _vp, _vp]),
    "thc_czt_half": (_i, [_vp, _vp, _vp, _l, _i, _i, _vp, _i, _vp]),
    "thc_isdf_half": (_i, [_vp, _vp, _vp, _l, _i, _i, _vp, _i, ctypes.POINTER(_i), _vp, ctypes.POINTER(IsdfOpts),
                           _vp, _vp, ctypes.POINTER(IsdfStats)]),
    "thc_czt_half_complex": (_i, [_vp, _vp, _vp, _l, _i, _i, _vp, _i, _vp]),
    "thc_isdf_half_complex": (_i, [_vp, _vp, _vp, _l, _i, _i, _vp, _i, ctypes.POINTER(_i), _vp,
                                   ctypes.POINTER(IsdfOpts), _vp, _vp, ctypes.POINTER(IsdfStats)]),
    "thc_factorize_batched": (_i, [_vp, _i, ctypes.POINTER(System), ctypes.POINTER(IsdfOpts), _i]),
}

_lib = None


class ThcError(RuntimeError):
    def __init__(self, status: int, where: str = ""):
        self.status = status
        msg = status_string(status)
        err = lib().thc_last_error().decode()
        super().__init__("%s: status %d (%s)%s" % (where, status, msg, (" -- " + err) if err else ""))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libthcpp_amd.so is not built (run: python -m thcpp_amd.build)")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def status_string(st: int) -> str:
    return lib().thc_status_string(st).decode()


def check(st: int, where: str, allow=()) -> int:
    if st != THC_OK and st not in allow:
        raise ThcError(st, where)
    return st
