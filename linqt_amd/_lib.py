"""ctypes binding of ``liblinqt_amd.so`` (the C ABI in ``include/linqt_amd.h``).

There is no CPU fallback: if the library or a GPU is missing, every entry
point raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblinqt_amd.so")

LQ_OK = 0
LQ_ERR_ARG = -1
LQ_ERR_HIP = -2
LQ_ERR_NOMEM = -3
LQ_ERR_UNSUPPORTED = -4
LQ_ERR_NODEV = -5

LQ_FILL_HERMITIAN_UPPER = 0
LQ_FILL_GENERAL = 1
LQ_KUBO_GREENWOOD = 0
LQ_KUBO_BASTIN = 1
LQ_ALGO_DIRECT = 0
LQ_ALGO_DOUBLING = 1
LQ_PREC_FP64 = 0
LQ_PREC_FP32X2 = 2
LQ_PREC_BF16X3 = 3
LQ_LIN_SPECTRAL, LQ_LIN_GREENWOOD, LQ_LIN_BASTIN, LQ_LIN_BASTIN_I, LQ_LIN_BASTIN_II, LQ_LIN_BASTIN_SEA, \
    LQ_LIN_BASTIN_SURF = range(7)
LQ_FMT_TEXT, LQ_FMT_TEXT_REF, LQ_FMT_BINARY = range(3)
LQ_PREC_INT8X4 = 4
LQ_PREC_INT8X5 = 5
LQ_PREC_INT8X6 = 6
LQ_PREC_INT8X7 = 7
LQ_PREC_INT8X8 = 8

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class lq_timing(C.Structure):
    _fields_ = [
        ("spmv_ms", C.c_double),
        ("contract_ms", C.c_double),
        ("post_ms", C.c_double),
        ("spmv_bytes", C.c_double),
        ("contract_flops", C.c_double),
        ("spmv_launches", C.c_longlong),
        ("contract_launches", C.c_longlong),
    ]


class lq_mom_header(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("M0", C.c_int),
        ("M1", C.c_int),
        ("n", C.c_longlong),
        ("W", C.c_double),
        ("Ec", C.c_double),
        ("dt", C.c_double),
    ]


# name -> (restype, argtypes); every function returns int unless noted
_SIGS = {
    "lq_last_error": (C.c_char_p, []),
    "lq_abi_version": (C.c_int, []),
    "lq_device_count": (C.c_int, [_ip]),
    "lq_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "lq_ctx_destroy": (C.c_int, [_vp]),
    "lq_ctx_synchronize": (C.c_int, [_vp]),
    "lq_ctx_release": (C.c_int, [_vp]),
    "lq_ctx_set_timing": (C.c_int, [_vp, C.c_int]),
    "lq_ctx_reset_timing": (C.c_int, [_vp]),
    "lq_ctx_get_timing": (C.c_int, [_vp, C.POINTER(lq_timing)]),
    "lq_malloc": (C.c_int, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "lq_free": (C.c_int, [_vp, _vp]),
    "lq_memcpy_h2d": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "lq_memcpy_d2h": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "lq_memcpy_d2d": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "lq_csr_create": (C.c_int, [_vp, C.c_int, C.c_int, _ip, _ip, _dp, C.c_int, C.POINTER(_vp)]),
    "lq_csr_destroy": (C.c_int, [_vp]),
    "lq_csr_rescale": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_double]),
    "lq_csr_info": (C.c_int, [_vp, _ip, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "lq_spmv": (C.c_int, [_vp, C.c_double, C.c_double, _vp, C.c_double, C.c_double, _vp, C.c_int]),
    "lq_compute_mom_table": (C.c_int, [_vp, _vp, C.c_longlong, C.c_int, _vp, C.c_longlong, C.c_int, C.c_longlong,
                                       _vp, C.c_int]),
    "lq_compute_mom_table_prec": (C.c_int, [_vp, _vp, C.c_longlong, C.c_int, _vp, C.c_longlong, C.c_int,
                                            C.c_longlong, C.c_int, _vp, C.c_int]),
    "lq_time_correlations": (C.c_int, [_vp, _vp, _vp, _dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                       C.c_double, _dp]),
    "lq_jackson_td": (C.c_int, [_vp, _dp, C.c_int, C.c_int, C.c_double, C.c_double, _ip]),
    "lq_timecorr_post": (C.c_int, [_vp, _dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                   C.c_double, _dp, _dp]),
    "lq_linear_grid": (C.c_int, [_vp, C.c_int, _dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                 _dp, _dp]),
    "lq_batch_create": (C.c_int, [_vp, C.c_size_t, C.c_int, C.POINTER(_vp)]),
    "lq_batch_destroy": (C.c_int, [_vp]),
    "lq_batch_job_bytes": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_size_t)]),
    "lq_batch_submit": (C.c_int, [_vp, _vp, _dp, C.c_int, C.c_int, _ip]),
    "lq_batch_run": (C.c_int, [_vp, C.c_int, _ip, _ip]),
    "lq_batch_poll": (C.c_int, [_vp, C.c_int, _ip]),
    "lq_batch_result": (C.c_int, [_vp, C.c_int, _dp]),
    "lq_batch_unroll": (C.c_int, [_vp, C.c_int, _ip]),
    "lq_overlap_solve": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_int, _ip, _dp]),
    "lq_spectral_moments_nonorth": (C.c_int, [_vp, _vp, _vp, _dp, C.c_int, C.c_int, C.c_double, C.c_int, _dp,
                                              C.POINTER(C.c_longlong)]),
    "lq_csr_export": (C.c_int, [_vp, C.POINTER(C.c_longlong), _ip, _ip, _dp]),
    "lq_csr_from_hoppings": (C.c_int, [_vp, C.c_int, C.c_int, _ip, _ip, _dp, _ip, C.POINTER(_vp)]),
    "lq_csr_file_read": (C.c_int, [C.c_char_p, _ip, C.POINTER(C.c_longlong), _ip, _ip, _dp]),
    "lq_csr_file_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_longlong, _ip, _ip, _dp]),
    "lq_csr_file_load": (C.c_int, [_vp, C.c_char_p, C.c_int, C.POINTER(_vp)]),
    "lq_mom_file_read": (C.c_int, [C.c_char_p, C.c_void_p, _dp]),
    "lq_mom_file_write": (C.c_int, [C.c_char_p, C.c_int, C.c_void_p, _dp]),
    "lq_spectral_moments": (C.c_int, [_vp, _vp, _dp, C.c_int, C.c_int, C.c_int, _dp]),
    "lq_spectral_moments_noalloc": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, _vp]),
    "lq_csr_tune": (C.c_int, [_vp, C.c_int, C.c_int, _ip, _ip, _ip, _ip]),
    "lq_hbm_probe": (C.c_int, [_vp, C.c_size_t, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "lq_correlation_moments": (C.c_int, [_vp, _vp, C.c_int, C.POINTER(_vp), _dp, C.c_int, C.c_int, C.c_int,
                                         C.c_int, C.c_int, _dp]),
    "lq_kubo_fft": (C.c_int, [_vp, _vp, C.c_int, C.POINTER(_vp), _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                              C.c_int, C.c_int, _dp]),
    "lq_kubo_post": (C.c_int, [_vp, C.c_int, _dp, C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp]),
    "lq_jackson_1d": (C.c_int, [_vp, _dp, C.c_int, C.c_double, C.c_double, _ip]),
    "lq_jackson_2d": (C.c_int, [_vp, _dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _ip, _ip]),
    "lq_dos_fftgrid": (C.c_int, [_vp, _dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp]),
    "lq_greenwood_fftgrid": (C.c_int, [_vp, _dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                       _dp, _dp]),
    "lq_bastin_fftgrid": (C.c_int, [_vp, _dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                    _dp, _dp]),
}

_lib = None


class LinqtError(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


def load():
    """Loads the library (no GPU call happens here)."""
    global _lib
    if _lib is None:
        if not os.path.isfile(LIB_PATH):
            raise RuntimeError(f"liblinqt_amd.so is not built ({LIB_PATH}); run `python -m linqt_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def call(name: str, *args) -> int:
    L = load()
    rc = getattr(L, name)(*args)
    if rc != LQ_OK:
        msg = L.lq_last_error()
        raise LinqtError(name, rc, msg.decode() if msg else "")
    return rc
