"""psfun_d_krylov: Krylov f(A)x method object (krylov/psfun_d_krylov_mod.F90).

`set` mirrors psfun_d_setstring/psfun_d_setinteger (lines 179-223);
`apply` mirrors psfun_d_parallel_apply (lines 225-269), dispatching on
psb_toupper(kname) to the device LANCZOS / ARNOLDI / SAIARNOLDI
implementations (SAIARNOLDI takes kmethd, prec, tau and the inner
itmax/itrace/istopc from the object, lines 262-264).  The
whole Krylov loop, the f(T) step and the stopping test run on the GPU.

apply(fun, a, desc_a, y, x, eps, itmax=None, itrace=None, istop=None)
    -> (info, iter, err, res)    with y overwritten in place.
"""
import ctypes as C

import numpy as np

from . import _lib
from . import psb


class psfun_d_krylov:
    def __init__(self):
        self.kname = "ARNOLDI"
        self.kmethd = "CG"
        self.prec = None          # inner preconditioner (shifted.psb_dprec_type) for SAIARNOLDI
        self.tau = 0.0
        self.itmax = 200
        self.itrace = 0
        self.istopc = 1
        self.irst = 10

    def set(self, what, val):
        key = str(what).upper()
        if isinstance(val, str):
            if key == "KNAME":
                self.kname = val
                return psb.psb_success_
            return psb.psb_err_invalid_args_combination_
        if key == "ITMAX":
            self.itmax = int(val)
        elif key == "ITRACE":
            self.itrace = int(val)
        elif key == "ISTOPC":
            self.istopc = int(val)
        elif key == "IRST":
            self.irst = int(val)
        else:
            return psb.psb_err_invalid_args_combination_
        return psb.psb_success_

    def apply(self, fun, a, desc_a, y, x, eps, itmax=None, itrace=None, istop=None):
        name = str(self.kname).upper()
        if name == "LANCZOS":
            return lanczos_apply(fun, a, desc_a, y, x, eps, itmax, itrace, istop)
        if name == "ARNOLDI":
            from .arnoldi import arnoldi_apply
            return arnoldi_apply(fun, a, desc_a, y, x, eps, itmax, itrace, istop)
        if name == "SAIARNOLDI":   # inner method, preconditioner, shift and inner limits from the object
            from .arnoldi import saiarnoldi_apply
            return saiarnoldi_apply(fun, a, self.kmethd, self.prec, self.tau, desc_a, self.itmax, self.itrace,
                                    self.istopc, y, x, eps, itmax, itrace, istop)
        return psb.psb_err_invalid_args_combination_, 0, 0.0, None


def lanczos_apply(fun, a, desc_a, y, x, eps, itmax=None, itrace=None, istop=None, opscale=1.0):
    """psfun_d_lanczos (krylov/polynomial/psfun_d_lanczos.f90:37-400).

    As in the reference, fun%scaling is NOT applied: the f(T) step uses
    scaling 1 and the Krylov operator is A (lines 172-173, 198, 239).
    `opscale` (default 1) is an extension for callers that want
    f(opscale*A)x without forming opscale*A."""
    istop_ = 1 if istop is None else int(istop)
    if istop_ < 1 or istop_ > 2:
        return psb.psb_err_invalid_istop_, 0, float(psb.psb_err_invalid_istop_), None
    nl = 200 if itmax is None else int(itmax)
    if x is None or x.v is None or y is None or y.v is None:
        return psb.psb_err_invalid_vect_state_, 0, 0.0, None
    n = desc_a.get_global_rows() if desc_a is not None else a.n
    if x.get_nrows() != n or y.get_nrows() != n or a.n != n:
        return psb.psb_err_from_subroutine_, 0, 0.0, None
    st, fn, _var = fun.device_fn()
    if st != _lib.PSFUN_SUCCESS:
        return psb.psb_err_from_subroutine_non_, 0, 0.0, None
    ctx = a.ctxt
    ctx.bind_stream()
    it = C.c_int(0)
    err = C.c_double(0.0)
    nres = C.c_int(0)
    res = np.zeros(max(nl, 2) + 2)
    with fun.bound(ctx):
        st = _lib.lib().psfun_lanczos_apply(ctx.handle, a.handle, fn, 1.0, float(opscale),
                                            C.c_void_p(x.v.data_ptr()), C.c_void_p(y.v.data_ptr()), float(eps), nl,
                                            istop_, C.byref(it), C.byref(err),
                                            res.ctypes.data_as(C.POINTER(C.c_double)), C.byref(nres))
    if st != _lib.PSFUN_SUCCESS:
        return psb.status_to_info(st), 0, 0.0, None
    psb.log_trace(f"LANCZOS+{fun.fname}+{fun.variant}", itrace, it.value, err.value, res[: nres.value], float(eps),
                  estimates=istop_ == 1)
    return psb.psb_success_, it.value, err.value, res[: nres.value].copy()
