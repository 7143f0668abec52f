"""ctypes binding of libadmixhip.so (the C ABI declared in include/admixhip.h).

The library is required: importing the package without it raises, and there
is no CPU fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libadmixhip.so")

ADMX_OK, ADMX_EINVAL, ADMX_EHIP, ADMX_ESTATE, ADMX_EIO, ADMX_EHOST = 0, 1, 2, 3, 4, 5
ADMX_F64, ADMX_F32 = 0, 1
ADMX_X, ADMX_X_NEXT, ADMX_X_NEXT2, ADMX_X_TMP = 0, 1, 2, 3
ADMX_SIDE_Q, ADMX_SIDE_P = 0, 1
ADMX_TUNE_TARGET_WAVES, ADMX_TUNE_KEEP_HESSIANS, ADMX_TUNE_PRECISION, ADMX_TUNE_GRAPHS = 0, 1, 2, 3
ADMX_TUNE_QN_MODE = 4
ADMX_TUNE_SKIPMISSING = 5
ADMX_TUNE_RESIDENT_SNPS = 6
ADMX_TUNE_SPLIT_Q, ADMX_TUNE_SPLIT_P = 7, 8
ADMX_TUNE_QSIDE_MFMA, ADMX_TUNE_PSIDE_MFMA = 9, 10
ADMX_QN_ZAL, ADMX_QN_LBQN = 0, 1
ADMX_PREC_F64, ADMX_PREC_F32, ADMX_PREC_BF16 = 0, 1, 2
ADMX_SKFR_GLOBAL, ADMX_SKFR_LOCAL = 0, 1
ADMX_EM_BASE, ADMX_EM_FALLBACK, ADMX_EM_FALLBACK2 = 0, 1, 2
ADMX_NOALLOC_LL_PARTS = 65536  # doubles of admx_loglikelihood_noalloc's fixed scratch

# Every exported symbol with its C signature (return type is int unless noted).
_i64, _int, _dbl, _vp, _u64 = C.c_int64, C.c_int, C.c_double, C.c_void_p, C.c_uint64
SIGNATURES = {
    "admx_version": ([], C.c_char_p),
    "admx_last_error": ([], C.c_char_p),
    "admx_device_count": ([_vp], _int),
    "admx_create": ([_vp, _i64, _i64, _i64, _i64, _int], _int),
    "admx_create_streamed": ([_vp, _i64, _i64, _i64, _i64, _int, _i64], _int),
    "admx_create_streamed_file": ([_vp, C.c_char_p, _i64, _i64, _i64, _i64, _int, _i64], _int),
    "admx_destroy": ([_vp], _int),
    "admx_upload_bed": ([_vp, _vp, _i64], _int),
    "admx_download_bed": ([_vp, _vp], _int),
    "admx_generate_bed": ([_vp, _vp, _vp, _u64, _dbl], _int),
    "admx_decode": ([_vp, _vp], _int),
    "admx_counts": ([_vp, _vp, _vp, _vp], _int),
    "admx_bed_from": ([_vp, _vp, C.c_int32, C.c_int32, C.c_uint64], _int),
    "admx_mlp_create": ([_vp, C.c_int32, C.c_int32, C.c_int32, C.c_uint64], _int),
    "admx_mlp_destroy": ([_vp], _int),
    "admx_mlp_num_params": ([_vp, _vp], _int),
    "admx_mlp_get_params": ([_vp, _vp], _int),
    "admx_mlp_set_params": ([_vp, _vp], _int),
    "admx_mlp_loss_grad": ([_vp, _vp, _vp, _i64, _vp, _vp], _int),
    "admx_mlp_predict": ([_vp, _vp, _i64, _vp], _int),
    "admx_mlp_train": ([_vp, _vp, _vp, _i64, C.c_int32, C.c_int32, C.c_float, C.c_uint64, _vp], _int),
    "admx_features": ([_vp, _vp, _vp, _vp, _i64, _vp], _int),
    "admx_warm_q": ([_vp, _vp], _int),
    "admx_warm_p": ([_vp, _vp], _int),
    "admx_copy_x": ([_vp, _int, _int], _int),
    "admx_heldout_deviance": ([_vp, _vp, C.c_int32, C.c_int32, C.c_uint64, _vp, _vp], _int),
    "admx_skfr_init": ([_vp, C.c_int32, C.c_uint64, C.c_int32, _vp], _int),
    "admx_skfr_run": ([_vp, C.c_int32, C.c_int32, _i64, C.c_int32, C.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
                      _int),
    "admx_loglikelihood_loop": ([_vp, _vp, _vp, _i64, _i64, _i64, _i64, _vp], _int),
    "admx_em_q_loop": ([_vp, _vp, _vp, _vp, _i64, _i64, _i64, _i64], _int),
    "admx_em_p_loop": ([_vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _i64], _int),
    "admx_update_q_loop": ([_vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _i64], _int),
    "admx_update_p_loop": ([_vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _i64], _int),
    "admx_qp_batch": ([_vp, _int, _i64, _vp, _vp, _vp, _vp, _vp], _int),
    "admx_project_q": ([_vp, _vp, _i64], _int),
    "admx_project_p": ([_vp, _vp, _i64], _int),
    "admx_lbqn_step": ([_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp], _int),
    "admx_zal_step": ([_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp], _int),
    "admx_set_x": ([_vp, _int, _vp], _int),
    "admx_get_x": ([_vp, _int, _vp], _int),
    "admx_set_uv": ([_vp, _vp, _vp], _int),
    "admx_get_uv": ([_vp, _vp, _vp], _int),
    "admx_set_ll": ([_vp, _dbl, _dbl], _int),
    "admx_get_ll": ([_vp, _vp, _vp], _int),
    "admx_get_hessians": ([_vp, _int, _vp, _vp], _int),
    "admx_loglikelihood": ([_vp, _int, _vp], _int),
    "admx_loglikelihood2": ([_vp, _int, _int, _vp], _int),
    "admx_em_q": ([_vp], _int),
    "admx_em_p": ([_vp], _int),
    "admx_init_em": ([_vp, _i64, _vp], _int),
    "admx_update_q": ([_vp, _int, _vp], _int),
    "admx_update_p": ([_vp, _int, _vp], _int),
    "admx_qn_iteration": ([_vp, _i64, _int, _int, _vp], _int),
    "admx_synchronize": ([_vp], _int),
    "admx_em_q_mode": ([_vp, _int], _int),
    "admx_em_p_mode": ([_vp, _int], _int),
    "admx_set_block": ([_vp, _int, _int, _vp], _int),
    "admx_get_block": ([_vp, _int, _int, _vp], _int),
    "admx_copy_block": ([_vp, _int, _int, _int], _int),
    "admx_grad_hessian": ([_vp, _int], _int),
    "admx_loglikelihood_noalloc": ([_vp, _i64, _i64, _i64, _i64, _vp, _vp, _vp, _vp], _int),
    "admx_profile_enable": ([_vp, _int], _int),
    "admx_profile_reset": ([_vp], _int),
    "admx_profile_report": ([_vp, C.c_char_p, C.c_size_t], _int),
    "admx_set_tuning": ([_vp, _int, _i64], _int),
}


class AdmixHIPError(RuntimeError):
    """A failed libadmixhip call (the message is admx_last_error())."""


_lib = None


def lib():
    """The loaded library; raises if it is missing (build it with __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AdmixHIPError(
                f"{LIB_PATH} not found: build the HIP library first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != ADMX_OK:
        msg = lib().admx_last_error().decode()
        raise AdmixHIPError(f"libadmixhip error {rc}: {msg}")


def ptr(a):
    """Raw pointer of a numpy array (None passes a NULL)."""
    if a is None:
        return None
    return C.c_void_p(a.ctypes.data)


def device_count() -> int:
    n = C.c_int(0)
    check(lib().admx_device_count(C.byref(n)))
    return n.value


def dtype_code(dtype) -> int:
    dt = np.dtype(dtype)
    if dt == np.float64:
        return ADMX_F64
    if dt == np.float32:
        return ADMX_F32
    raise ValueError(f"unsupported element type {dt}; use Float64 or Float32")
