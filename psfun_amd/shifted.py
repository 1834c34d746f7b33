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
    """Preconditioner object with the kinds the shifted CG applies on the
    device, built on the matrix given to build() (the reference's drivers
    build it on the unshifted A, shiftedtest.f90:219-223):
      'NONE' / 'NOPREC'
      'DIAG'  the inverse diagonal
      'BJAC'  block Jacobi, which with one process is the whole matrix
              factored by ILU(0) (host factorization, device triangular
              solves by level sets; psfun_prec_create)."""

    def __init__(self):
        self.ptype = None
        self.built = False
        self.handle = None
        self.n = None

    def init(self, ctxt, ptype):
        p = str(ptype).upper()
        if p not in ("NONE", "NOPREC", "DIAG", "BJAC"):
            return psb.psb_err_invalid_args_combination_
        self.free()
        self.ptype = {"NONE": "NONE", "NOPREC": "NONE"}.get(p, p)
        self.built = False
        return psb.psb_success_

    def build(self, a, desc_a=None):
        if self.ptype is None:
            return psb.psb_err_invalid_args_combination_
        self.free()
        h = C.c_void_p()
        st = _lib.lib().psfun_prec_create(a.ctxt.handle, a.handle, self.code(), C.byref(h))
        if st != _lib.PSFUN_SUCCESS:
            return psb.psb_err_from_subroutine_
        self.handle = h
        self.n = a.n
        self.built = True
        return psb.psb_success_

    def code(self):
        return {"DIAG": _lib.PSFUN_PREC_DIAG, "BJAC": _lib.PSFUN_PREC_BJAC}.get(self.ptype, _lib.PSFUN_PREC_NONE)

    def levels(self):
        """Level-scheduled steps of one BJAC application (0 otherwise)."""
        if self.handle is None or self.ptype != "BJAC":
            return 0
        k, lv = C.c_int(), C.c_int()
        _lib.check(_lib.lib().psfun_prec_info(self.handle, C.byref(k), C.byref(lv)))
        return lv.value

    def apply(self, ctx, r, z):
        """z = M^{-1} r on device tensors (BJAC objects)."""
        ctx.bind_stream()
        _lib.check(_lib.lib().psfun_prec_apply(ctx.handle, self.handle, C.c_void_p(r.data_ptr()),
                                               C.c_void_p(z.data_ptr())), "psfun_prec_apply")

    def free(self):
        if getattr(self, "handle", None) is not None:
            _lib.lib().psfun_prec_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


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
    if prec is not None and getattr(prec, "handle", None) is not None:   # a built object
        st = _lib.lib().psfun_cg_shifted_prec(ctx.handle, a.handle, prec.handle, C.c_void_p(b.v.data_ptr()),
                                              float(eta), float(zeta), C.c_void_p(x.v.data_ptr()), float(eps),
                                              itmax_, istop_, C.byref(it), C.byref(err),
                                              C.byref(cnd) if want_cond else None)
    else:
        st = _lib.lib().psfun_cg_shifted(ctx.handle, a.handle, pc, C.c_void_p(b.v.data_ptr()), float(eta),
                                         float(zeta), C.c_void_p(x.v.data_ptr()), float(eps), itmax_, istop_,
                                         C.byref(it), C.byref(err), C.byref(cnd) if want_cond else None)
    if st != _lib.PSFUN_SUCCESS:
        return psb.psb_err_from_subroutine_, 0, 0.0, 0.0
    # PSBLAS's check_conv logs every itrace-th iteration inside the loop; the
    # device loop keeps no per-iteration history, so the log has its first
    # and last lines (the header and the final error)
    psb.log_trace("CG", itrace, it.value, err.value, [], float(eps), estimates=False)
    return psb.psb_success_, it.value, err.value, (cnd.value if want_cond else 0.0)


def mscg(a, b, shifts, eps=1e-8, itmax=1000, prec=None):
    """Solve (A + sigma_s I) X[s] = b for every complex shift on the GPU.
    With a built preconditioner object, the shifts the multi-shift solve
    leaves above eps after itmax iterations continue by preconditioned COCG
    (psfun_mscg_prec).  Returns (X as a complex128 torch tensor of shape
    (S, n), iters, resid)."""
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
    if prec is not None:
        if getattr(prec, "handle", None) is None:
            raise ValueError("mscg: the preconditioner is not built (prec.build(a))")
        st = _lib.lib().psfun_mscg_prec(ctx.handle, a.handle, prec.handle, S, sh.ctypes.data_as(C.c_void_p),
                                        C.c_void_p(b.v.data_ptr()), C.c_void_p(X.data_ptr()), n, float(eps),
                                        int(itmax), iters.ctypes.data_as(C.c_void_p),
                                        resid.ctypes.data_as(C.c_void_p))
        _lib.check(st, "psfun_mscg_prec")
    else:
        st = _lib.lib().psfun_mscg(ctx.handle, a.handle, S, sh.ctypes.data_as(C.c_void_p),
                                   C.c_void_p(b.v.data_ptr()), C.c_void_p(X.data_ptr()), n, float(eps), int(itmax),
                                   iters.ctypes.data_as(C.c_void_p), resid.ctypes.data_as(C.c_void_p))
        _lib.check(st, "psfun_mscg")
    return X, iters, resid
