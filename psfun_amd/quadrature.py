"""Rational-quadrature f(A)x: psfun_z_quadrature
(quadratures/psfun_z_quadrature_mod.F90:42-166) with its pole rules, and the
apply that the reference documents (docs/pages/algorithms.md:106-125) but does
not implement.

  quad = psfun_z_quadrature()
  info = quad.computepoles(hhtmethod1, zfun, N, rparams=[lmin, lmax])
  quad.set(a)                       # setmatrix
  info, iters, resid = quad.apply(x, y, eps, itmax)

Convention of the stored rule (xi, c, eta, sign): the shifted systems are
(xi_j I + sign*A) v_j = x.  The assembly depends on the rule:
  hhtmethod1 (HHT 2008 method 1):  y = eta * A * Im( sum_j c_j v_j )
  talbot (exp(-t z)):              y = eta * Re( sum_j c_j v_j )
Every shifted system is solved by one multi-shift CG on the GPU.  For a
resolvent-sum rule whose poles and weights come in complex-conjugate pairs
(Talbot), each pair needs only one solve: for real A and x the pair sums to
2 Re(c v).
"""
import ctypes as C

import numpy as np

from . import _lib
from . import psb

FORM_HHT1 = "HHT1"
FORM_RESOLVENT = "RESOLVENT"


def horner(coeffs, x):
    c = np.ascontiguousarray(coeffs, dtype=np.float64)
    return _lib.lib().psfun_horner(len(c), c.ctypes.data_as(C.c_void_p), float(x))


def ellipkkp(L):
    K = np.zeros(2)
    _lib.check(_lib.lib().psfun_ellipkkp(float(L), K.ctypes.data_as(C.c_void_p)), "ellipkkp")
    return K


def ellipj(u, L):
    out = np.zeros(6)
    u = complex(u)
    _lib.check(_lib.lib().psfun_ellipj(u.real, u.imag, float(L), out.ctypes.data_as(C.c_void_p)), "ellipj")
    sn, cn, dn = complex(out[0], out[1]), complex(out[2], out[3]), complex(out[4], out[5])
    if isinstance(u, complex) and u.imag == 0.0 and sn.imag == 0.0:
        return sn.real, cn.real, dn.real
    return sn, cn, dn


def _pairs_to_complex(a):
    return a[0::2] + 1j * a[1::2]


def hhtmethod1(zfun, N, cparams=None, rparams=None):
    """Returns (xi, c, eta, sign, info) (quadratures/psfun_z_quadrules_mod.f90:40-85)."""
    if rparams is None or len(rparams) < 2:
        return None, None, 0.0, 0.0, psb.psb_err_invalid_args_combination_

    def cb(z, fz, _user):
        v = complex(zfun(complex(z[0], z[1])))
        fz[0] = v.real
        fz[1] = v.imag

    cfun = _lib.ZFUN(cb)
    xi = np.zeros(2 * N)
    c = np.zeros(2 * N)
    eta = C.c_double(0.0)
    sign = C.c_double(0.0)
    rp = np.ascontiguousarray(rparams[:2], dtype=np.float64)
    st = _lib.lib().psfun_hhtmethod1(cfun, None, int(N), rp.ctypes.data_as(C.c_void_p), xi.ctypes.data_as(C.c_void_p),
                                     c.ctypes.data_as(C.c_void_p), C.byref(eta), C.byref(sign))
    if st != _lib.PSFUN_SUCCESS:
        return None, None, 0.0, 0.0, psb.psb_err_from_subroutine_
    return _pairs_to_complex(xi), _pairs_to_complex(c), eta.value, sign.value, psb.psb_success_


hhtmethod1.form = FORM_HHT1


def talbot(zfun, N, cparams=None, rparams=None):
    """Modified Talbot rule for exp(-t z); rparams = [t] (default t = 1).
    `zfun` is accepted for interface compatibility and must be exp(-z) (or None)."""
    t = 1.0 if rparams is None or len(rparams) == 0 else float(rparams[0])
    xi = np.zeros(2 * N)
    c = np.zeros(2 * N)
    eta = C.c_double(0.0)
    sign = C.c_double(0.0)
    st = _lib.lib().psfun_talbot(int(N), t, xi.ctypes.data_as(C.c_void_p), c.ctypes.data_as(C.c_void_p),
                                 C.byref(eta), C.byref(sign))
    if st != _lib.PSFUN_SUCCESS:
        return None, None, 0.0, 0.0, psb.psb_err_from_subroutine_
    return _pairs_to_complex(xi), _pairs_to_complex(c), eta.value, sign.value, psb.psb_success_


talbot.form = FORM_RESOLVENT


def fold_conjugate_pairs(shifts, weights, tol=1e-13):
    """Merge (s, w), (conj s, conj w) pairs into (s, 2w) for a Re(...) assembly."""
    shifts = np.asarray(shifts, np.complex128)
    weights = np.asarray(weights, np.complex128)
    used = np.zeros(len(shifts), bool)
    out_s, out_w = [], []
    for i in range(len(shifts)):
        if used[i]:
            continue
        used[i] = True
        s, w = shifts[i], weights[i]
        scale = max(abs(s), 1e-300)
        if abs(s.imag) > tol * scale:
            d = np.abs(shifts - np.conj(s)) + np.abs(weights - np.conj(w)) / max(abs(w), 1e-300) * scale
            d[used] = np.inf
            j = int(np.argmin(d))
            if d[j] <= tol * scale * 10:
                used[j] = True
                out_s.append(s)
                out_w.append(2.0 * w)
                continue
        out_s.append(s)
        out_w.append(w)
    return np.array(out_s), np.array(out_w)


class psfun_z_quadrature:
    def __init__(self):
        self.xi = None
        self.c = None
        self.eta = 0.0
        self.sign = 0.0
        self.a = None
        self.prec = None
        self.form = FORM_RESOLVENT

    def computepoles(self, quadformula, zfun, N, cparams=None, rparams=None):
        """psfun_z_computepoles (psfun_z_quadrature_mod.F90:114-142).  As in
        the reference, when cparams is present only cparams is passed."""
        if cparams is not None:
            xi, c, eta, sign, info = quadformula(zfun, N, cparams=cparams)
        elif rparams is not None:
            xi, c, eta, sign, info = quadformula(zfun, N, rparams=rparams)
        else:
            xi, c, eta, sign, info = quadformula(zfun, N)
        if info == psb.psb_success_:
            self.xi, self.c, self.eta, self.sign = xi, c, eta, sign
            self.form = getattr(quadformula, "form", FORM_RESOLVENT)
        return info

    def set(self, obj):
        """setmatrix / setpreconditioner (psfun_z_quadrature_mod.F90:144-166)."""
        if isinstance(obj, psb.psb_dspmat_type):
            self.a = obj
        else:
            self.prec = obj

    def apply(self, x, y, eps=1e-8, itmax=1000, fold_pairs=True):
        """y = f(A) x with the stored rule; returns (info, iters, resid).
        With a preconditioner set (set(prec); built on the quadrature's
        matrix here if it is not yet built), the poles the multi-shift solve
        leaves above eps after itmax iterations continue by preconditioned
        COCG (psfun_quad_apply_prec)."""
        if self.a is None or self.xi is None:
            return psb.psb_err_invalid_args_combination_, None, None
        if x is None or x.v is None or y is None or y.v is None:
            return psb.psb_err_invalid_vect_state_, None, None
        sg = float(self.sign)
        shifts = sg * np.asarray(self.xi, np.complex128)   # (xi I + sign A) = sign (A + sign xi I)
        weights = np.asarray(self.c, np.complex128)
        gamma = float(self.eta) * sg
        if self.form == FORM_HHT1:
            form = _lib.PSFUN_QUAD_A_IM_SUM
        else:
            form = _lib.PSFUN_QUAD_RE_SUM
            if fold_pairs:
                shifts, weights = fold_conjugate_pairs(shifts, weights)
        S = len(shifts)
        sh = np.ascontiguousarray(np.stack([shifts.real, shifts.imag], 1).reshape(-1))
        ww = np.ascontiguousarray(np.stack([weights.real, weights.imag], 1).reshape(-1))
        iters = np.zeros(S, dtype=np.int32)
        resid = np.zeros(S)
        ctx = self.a.ctxt
        ctx.bind_stream()
        if self.prec is not None:
            if getattr(self.prec, "handle", None) is None:
                info = self.prec.build(self.a)
                if info != psb.psb_success_:
                    return info, None, None
            st = _lib.lib().psfun_quad_apply_prec(ctx.handle, self.a.handle, self.prec.handle, S,
                                                  sh.ctypes.data_as(C.c_void_p), ww.ctypes.data_as(C.c_void_p),
                                                  gamma, form, C.c_void_p(x.v.data_ptr()), C.c_void_p(y.v.data_ptr()),
                                                  float(eps), int(itmax), iters.ctypes.data_as(C.c_void_p),
                                                  resid.ctypes.data_as(C.c_void_p))
        else:
            st = _lib.lib().psfun_quad_apply(ctx.handle, self.a.handle, S, sh.ctypes.data_as(C.c_void_p),
                                             ww.ctypes.data_as(C.c_void_p), gamma, form, C.c_void_p(x.v.data_ptr()),
                                             C.c_void_p(y.v.data_ptr()), float(eps), int(itmax),
                                             iters.ctypes.data_as(C.c_void_p), resid.ctypes.data_as(C.c_void_p))
        if st != _lib.PSFUN_SUCCESS:
            return psb.status_to_info(st), None, None
        return psb.psb_success_, iters, resid
