"""ctypes binding of the C-ABI in include/psfun_hip.h (libpsfun_hip.so).

The shared library is built in-tree by psfun_amd/build.py.  Loading fails
loudly when it is missing: there is no CPU fallback behind this package.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsfun_hip.so")

# status codes (include/psfun_hip.h)
PSFUN_SUCCESS = 0
PSFUN_ERR_INVALID_ARG = 1
PSFUN_ERR_INVALID_ISTOP = 2
PSFUN_ERR_INVALID_VECT_STATE = 3
PSFUN_ERR_INVALID_ARGS_COMBINATION = 4
PSFUN_ERR_NOT_SQUARE = 5
PSFUN_ERR_UNSUPPORTED = 6
PSFUN_ERR_ALLOC = 7
PSFUN_ERR_HIP = 8
PSFUN_ERR_NO_DEVICE = 9
PSFUN_ERR_BREAKDOWN = 10
PSFUN_ERR_TOO_LARGE = 11

PSFUN_FN_EXP = 0
PSFUN_FN_USERF = 1
VARIANTS = {"CHBSYM": 0, "CHBHES": 1, "CHBGEN": 2, "GENPADE": 3, "TAYLOR": 4, "SASQ": 5}
PSFUN_VAR_SYM = 6
PSFUN_MEM_HOST = 0
PSFUN_MEM_DEVICE = 1
PSFUN_FORMAT_AUTO = 0
PSFUN_FORMAT_CSR = 1
PSFUN_FORMAT_SELL = 2
PSFUN_FORMAT_HDIA = 3
PSFUN_FORMAT_CSRS = 4
PSFUN_PREC_NONE = 0
PSFUN_PREC_DIAG = 1
PSFUN_PREC_BJAC = 2
PSFUN_ORDER_NATURAL = 0
PSFUN_ORDER_RCM = 1
PSFUN_QUAD_RE_SUM = 0
PSFUN_QUAD_A_IM_SUM = 1

PROF_IDS = {"lz_spmv": 0, "lz_update": 1, "lz_assemble": 2, "spmv": 3, "ms_spmv": 4, "ms_update": 5,
            "cg": 6, "batch": 7, "sddmm": 8, "ar_spmv": 9, "ar_mgs": 10,
            "reduce": 11, "axpby": 12}

_vp = C.c_void_p
_i = C.c_int
_i64 = C.c_int64
_d = C.c_double
_pi = C.POINTER(C.c_int)
_pd = C.POINTER(C.c_double)
_pi64 = C.POINTER(C.c_int64)

# name -> (restype, argtypes); every function returns an int status unless noted
SIGNATURES = {
    "psfun_strerror": (C.c_char_p, [_i]),
    "psfun_ctx_create": (_i, [_i, C.POINTER(_vp)]),
    "psfun_ctx_destroy": (_i, [_vp]),
    "psfun_ctx_set_stream": (_i, [_vp, _vp]),
    "psfun_ctx_use_own_stream": (_i, [_vp]),
    "psfun_set_userf": (_i, [_vp, _vp, _vp]),
    "psfun_ctx_synchronize": (_i, [_vp]),
    "psfun_ctx_num_cus": (_i, [_vp, _pi]),
    "psfun_prof_enable": (_i, [_vp, _i]),
    "psfun_prof_reset": (_i, [_vp]),
    "psfun_prof_read": (_i, [_vp, _i, _pd, _pi64, _pd]),
    "psfun_csr_create": (_i, [_vp, _i64, _i64, _vp, _vp, _vp, _i, _i, C.POINTER(_vp)]),
    "psfun_csr_destroy": (_i, [_vp]),
    "psfun_csr_set_values": (_i, [_vp, _vp]),
    "psfun_csr_set_lanes": (_i, [_vp, _i]),
    "psfun_csr_info": (_i, [_vp, _pi64, _pi64, _pi]),
    "psfun_csr_set_format": (_i, [_vp, _i]),
    "psfun_csr_get_format": (_i, [_vp, _pi, _pi64]),
    "psfun_csr_sell_index_bits": (_i, [_vp, _pi]),
    "psfun_csr_hdia_diagonals": (_i, [_vp, _pi64]),
    "psfun_vec_alloc": (_i, [_vp, _i64, C.POINTER(_vp)]),
    "psfun_vec_free": (_i, [_vp, _vp]),
    "psfun_vec_upload": (_i, [_vp, _vp, _vp, _i64]),
    "psfun_vec_download": (_i, [_vp, _vp, _vp, _i64]),
    "psfun_spmv": (_i, [_vp, _vp, _d, _vp, _d, _vp]),
    "psfun_spmm": (_i, [_vp, _vp, _i, _d, _vp, _i64, _d, _vp, _i64]),
    "psfun_dot": (_i, [_vp, _i64, _vp, _vp, _pd]),
    "psfun_nrm2": (_i, [_vp, _i64, _vp, _pd]),
    "psfun_amax": (_i, [_vp, _i64, _vp, _pd]),
    "psfun_axpby": (_i, [_vp, _i64, _d, _vp, _d, _vp]),
    "psfun_scal": (_i, [_vp, _i64, _d, _vp]),
    "psfun_serial_apply_host": (_i, [_vp, _i, _i, _d, _i, _vp, _vp, _vp]),
    "psfun_lanczos_apply": (_i, [_vp, _vp, _i, _d, _d, _vp, _vp, _d, _i, _i, _pi, _pd, _pd, _pi]),
    "psfun_lanczos_apply_noalloc": (_i, [_vp, _vp, _i, _d, _d, _vp, _vp, _d, _i, _i, _pi, _pd, _pd, _pi]),
    "psfun_arnoldi_apply": (_i, [_vp, _vp, _i, _i, _d, _d, _vp, _vp, _d, _i, _i, _pi, _pd, _pd, _pi]),
    "psfun_sai_arnoldi_apply": (_i, [_vp, _vp, _i, _i, _d, _d, _i, _i, _i, _vp, _vp, _d, _i, _i, _pi, _pd, _pd,
                                     _pi]),
    "psfun_batch_create": (_i, [_vp, _i, _vp, _vp, _vp, _vp, C.POINTER(_vp)]),
    "psfun_batch_create_ordered": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _i, C.POINTER(_vp)]),
    "psfun_rcm_order": (_i, [_i, _vp, _vp, _i, _vp]),
    "psfun_batch_destroy": (_i, [_vp]),
    "psfun_batch_set_tuning": (_i, [_vp, _i, _i]),
    "psfun_batch_info": (_i, [_vp, _pi, _pi64, _pi64, _pi, _pi]),
    "psfun_batch_lanczos": (_i, [_vp, _vp, _i, _d, _d, _vp, _vp, _d, _i, _i, _vp, _vp]),
    "psfun_batch_spmm": (_i, [_vp, _vp, _i, _d, _vp, _d, _vp]),
    "psfun_batch_set_values": (_i, [_vp, _vp]),
    "psfun_batch_csr": (_i, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "psfun_sddmm_shifted": (_i, [_vp, _i, _vp, _vp, _i, _vp, _vp, _i64, _vp, _i64, _d, _vp]),
    "psfun_batch_lanczos_basis": (_i, [_vp, _vp, _i, _d, _d, _vp, _vp, _d, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    "psfun_krylov_resolvents": (_i, [_vp, _i64, _i, _i, _vp, _vp, _vp, _i, _d, _vp, _vp, _i, _vp, _vp]),
    "psfun_sddmm_cols": (_i, [_vp, _i, _vp, _vp, _i, _i, _vp, _vp, _vp, _d, _vp]),
    "psfun_expv_vjp": (_i, [_vp, _vp, _i, _d, _vp, _vp, _vp, _d, _i, _i, _vp, _pd, _vp]),
    "psfun_cg_shifted": (_i, [_vp, _vp, _i, _vp, _d, _d, _vp, _d, _i, _i, _pi, _pd, _pd]),
    "psfun_set_pade_degree": (_i, [_vp, _i]),
    "psfun_prec_create": (_i, [_vp, _vp, _i, C.POINTER(_vp)]),
    "psfun_prec_destroy": (_i, [_vp]),
    "psfun_prec_info": (_i, [_vp, _pi, _pi]),
    "psfun_prec_apply": (_i, [_vp, _vp, _vp, _vp]),
    "psfun_cg_shifted_prec": (_i, [_vp, _vp, _vp, _vp, _d, _d, _vp, _d, _i, _i, _pi, _pd, _pd]),
    "psfun_mscg_prec": (_i, [_vp, _vp, _vp, _i, _vp, _vp, _vp, _i64, _d, _i, _vp, _vp]),
    "psfun_quad_apply_prec": (_i, [_vp, _vp, _vp, _i, _vp, _vp, _d, _i, _vp, _vp, _d, _i, _vp, _vp]),
    "psfun_mscg": (_i, [_vp, _vp, _i, _vp, _vp, _vp, _i64, _d, _i, _vp, _vp]),
    "psfun_horner": (_d, [_i, _vp, _d]),
    "psfun_ellipkkp": (_i, [_d, _vp]),
    "psfun_ellipj": (_i, [_d, _d, _d, _vp]),
    "psfun_hhtmethod1": (_i, [_vp, _vp, _i, _vp, _vp, _vp, _pd, _pd]),
    "psfun_talbot": (_i, [_i, _d, _vp, _vp, _pd, _pd]),
    "psfun_quad_apply": (_i, [_vp, _vp, _i, _vp, _vp, _d, _i, _vp, _vp, _d, _i, _vp, _vp]),
}

ZFUN = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p)
USERF = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)   # psfun_userf_t

_lib = None


class PsfunError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = lib().psfun_strerror(status).decode() if _lib is not None else str(status)
        super().__init__(f"{what}: {msg} (status {status})" if what else f"{msg} (status {status})")


def lib():
    """Load libpsfun_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m psfun_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status, what=""):
    if status != PSFUN_SUCCESS:
        raise PsfunError(status, what)
    return status


def exported_symbols():
    return list(SIGNATURES)
