"""psfun_d_arnoldi (krylov/polynomial/psfun_d_arnoldi.f90:37-352) on the GPU.

The whole loop runs on the device: SpMV, modified Gram-Schmidt, the f(H) step
with the reference's a-posteriori estimate, and the assembly.  As in the
reference, the SpMV is multiplied by fun%scaling while the f(H) step uses
scaling 1 (lines 172-173, 202); f(H) reads H as fun%variant reads a matrix.

arnoldi_apply(fun, a, desc_a, y, x, eps, itmax=None, itrace=None, istop=None)
    -> (info, iter, err, res)    with y overwritten in place.
"""
import ctypes as C

import numpy as np

from . import _lib
from . import psb


def arnoldi_apply(fun, a, desc_a, y, x, eps, itmax=None, itrace=None, istop=None):
    istop_ = 1 if istop is None else int(istop)            # arnoldi:99-110
    if istop_ < 1 or istop_ > 2:
        return psb.psb_err_invalid_istop_, 0, float(psb.psb_err_invalid_istop_), None
    nl = 200 if itmax is None else int(itmax)              # arnoldi:113-118
    if x is None or x.v is None or y is None or y.v is None:
        return psb.psb_err_invalid_vect_state_, 0, 0.0, None
    n = desc_a.get_global_rows() if desc_a is not None else a.n
    if x.get_nrows() != n or y.get_nrows() != n or a.n != n:
        return psb.psb_err_from_subroutine_, 0, 0.0, None
    st, fn, var = fun.device_fn()
    if st != _lib.PSFUN_SUCCESS:
        return psb.psb_err_from_subroutine_non_, 0, 0.0, None
    ctx = a.ctxt
    ctx.bind_stream()
    it = C.c_int(0)
    err = C.c_double(0.0)
    nres = C.c_int(0)
    res = np.zeros(max(nl, 2) + 2)
    with fun.bound(ctx):
        st = _lib.lib().psfun_arnoldi_apply(ctx.handle, a.handle, fn, var, 1.0, float(fun.scaling),
                                            C.c_void_p(x.v.data_ptr()), C.c_void_p(y.v.data_ptr()), float(eps), nl,
                                            istop_, C.byref(it), C.byref(err),
                                            res.ctypes.data_as(C.POINTER(C.c_double)), C.byref(nres))
    if st != _lib.PSFUN_SUCCESS:
        return psb.status_to_info(st), 0, 0.0, None
    psb.log_trace(f"ARNOLDI+{fun.fname}+{fun.variant}", itrace, it.value, err.value, res[: nres.value], float(eps),
                  estimates=istop_ == 1)
    return psb.psb_success_, it.value, err.value, res[: nres.value].copy()


def saiarnoldi_apply(fun, a, kryl, prec, tau, desc_a, initmax, initrace, inistop, y, x, eps, itmax=None,
                     itrace=None, istop=None):
    """psfun_d_saiarnoldi (krylov/rational/psfun_d_sai.f90:37-419) on the GPU:
    Arnoldi on (A + tau I)^{-1}, every inner solve by the device PCG (`kryl`
    'CG'; any other name warns and runs CG, as psfun_dkrylov_vect does;
    `prec` a psb_dprec_type or None), then f(H_k^{-1} - tau I).  As in the
    reference fun%scaling is not applied (lines 172-173)."""
    import sys
    istop_ = 1 if istop is None else int(istop)
    if istop_ < 1 or istop_ > 2:
        return psb.psb_err_invalid_istop_, 0, float(psb.psb_err_invalid_istop_), None
    nl = 200 if itmax is None else int(itmax)
    if x is None or x.v is None or y is None or y.v is None:
        return psb.psb_err_invalid_vect_state_, 0, 0.0, None
    n = desc_a.get_global_rows() if desc_a is not None else a.n
    if x.get_nrows() != n or y.get_nrows() != n or a.n != n:
        return psb.psb_err_from_subroutine_, 0, 0.0, None
    if str(kryl).upper() != "CG":
        sys.stderr.write(f"psfun_krylov: Warning: Unknown method  {kryl}, defaulting to CG\n")
    st, fn, var = fun.device_fn()
    if st != _lib.PSFUN_SUCCESS:
        return psb.psb_err_from_subroutine_non_, 0, 0.0, None
    pc = _lib.PSFUN_PREC_NONE if prec is None else prec.code()
    ctx = a.ctxt
    ctx.bind_stream()
    it = C.c_int(0)
    err = C.c_double(0.0)
    nres = C.c_int(0)
    res = np.zeros(max(nl, 2) + 2)
    with fun.bound(ctx):
        st = _lib.lib().psfun_sai_arnoldi_apply(ctx.handle, a.handle, fn, var, 1.0, float(tau), pc,
                                                int(initmax), int(inistop), C.c_void_p(x.v.data_ptr()),
                                                C.c_void_p(y.v.data_ptr()), float(eps), nl, istop_, C.byref(it),
                                                C.byref(err), res.ctypes.data_as(C.POINTER(C.c_double)),
                                                C.byref(nres))
    if st != _lib.PSFUN_SUCCESS:
        return psb.status_to_info(st), 0, 0.0, None
    psb.log_trace(f"SAIARNOLDI+{fun.fname}+{fun.variant}", itrace, it.value, err.value, res[: nres.value], float(eps),
                  estimates=istop_ == 1)
    return psb.psb_success_, it.value, err.value, res[: nres.value].copy()

