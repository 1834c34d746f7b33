"""Shifted linear systems: the psb_krylov extension of PSFUN
(krylov/psfun_krylov_mod.f90:39-142) over the device CG of
krylov/shifted/psfun_dcg.F90, plus the multi-shift solver used by the
quadrature apply.

psb_krylov(method, a, prec, b, eta, zeta, x, eps, desc_a, itmax=None,
           itrace=None, irst=None, istop=None, cond=False)
    -> (info, iter, err, cond)      solves (eta*A + zeta*I) x = b in place.
"""
import ctypes as C
import sys

import numpy as np

from . import _lib
from . import psb


class psb_dprec_type:
    """Preconditioner object with the two kinds the shifted CG supports on
    the device: 'NONE' and 'DIAG' (diagonal of the matrix it was built on;
    the reference's driver builds it on the unshifted A)."""

    def __init__(self):
        self.ptype = None
        self.built = False

    def init(self, ctxt, ptype):
        p = str(ptype).upper()
        if p not in ("NONE", "NOPREC", "DIAG"):
            return psb.psb_err_invalid_args_combination_
        self.ptype = "NONE" if p in ("NONE", "NOPREC") else "DIAG"
        return psb.psb_success_

    def build(self, a, desc_a):
        if self.ptype is None:
            return psb.psb_err_invalid_args_combination_
        self.built = True
        return psb.psb_success_

    def code(self):
        return _lib.PSFUN_PREC_DIAG if self.ptype == "DIAG" else _lib.PSFUN_PREC_NONE


def psb_krylov(method, a, prec, b, eta, zeta, x, eps, desc_a, itmax=None, itrace=None, irst=None, istop=None,
               cond=False):
    """psfun_dkrylov_vect: only 'CG' is implemented; any other name warns and
    runs CG without the condition estimate (psfun_krylov_mod.f90:116-126)."""
    m = str(method).upper()
    want_cond = bool(cond)
    if m != "CG":
        sys.stderr.write(f"psfun_krylov: Warning: Unknown method  {method}, defaulting to CG\n")
        want_cond = False
    if b is None or b.v is None or x is None or x.v is None:
        return psb.psb_err_invalid_vect_state_, 0, 0.0, 0.0
    istop_ = 2 if istop is None else int(istop)          # psfun_dcg.F90:193-197
    itmax_ = 1000 if itmax is None else int(itmax)       # psfun_dcg.F90:225-229
    if istop_ not in (1, 2):
        return psb.psb_err_invalid_istop_, 0, 0.0, 0.0
    pc = _lib.PSFUN_PREC_NONE if prec is None else prec.code()
    ctx = a.ctxt
    ctx.bind_stream()
    it = C.c_int(0)
    err = C.c_double(0.0)
    cnd = C.c_double(0.0)
    st = _lib.lib().psfun_cg_shifted(ctx.handle, a.handle, pc, C.c_void_p(b.v.data_ptr()), float(eta), float(zeta),
                                     C.c_void_p(x.v.data_ptr()), float(eps), itmax_, istop_, C.byref(it),
                                     C.byref(err), C.byref(cnd) if want_cond else None)
    if st != _lib.PSFUN_SUCCESS:
        return psb.psb_err_from_subroutine_, 0, 0.0, 0.0
    return psb.psb_success_, it.value, err.value, (cnd.value if want_cond else 0.0)


def mscg(a, b, shifts, eps=1e-8, itmax=1000):
    """Solve (A + sigma_s I) X[s] = b for every complex shift on the GPU.
    Returns (X as a complex128 torch tensor of shape (S, n), iters, resid)."""
    import torch
    shifts = np.asarray(shifts, dtype=np.complex128).reshape(-1)
    S = len(shifts)
    n = a.n
    sh = np.ascontiguousarray(np.stack([shifts.real, shifts.imag], 1).reshape(-1))
    X = torch.zeros((S, n), dtype=torch.complex128, device=b.v.device)
    iters = np.zeros(S, dtype=np.int32)
    resid = np.zeros(S)
    ctx = a.ctxt
    ctx.bind_stream()
    st = _lib.lib().psfun_mscg(ctx.handle, a.handle, S, sh.ctypes.data_as(C.c_void_p), C.c_void_p(b.v.data_ptr()),
                               C.c_void_p(X.data_ptr()), n, float(eps), int(itmax), iters.ctypes.data_as(C.c_void_p),
                               resid.ctypes.data_as(C.c_void_p))
    _lib.check(st, "psfun_mscg")
    return X, iters, resid
