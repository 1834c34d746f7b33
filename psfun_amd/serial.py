"""psfun_d_serial: the function object of the f(H) step
(serial/psfun_d_serial_mod.F90:43-170).

Holds fname ('EXP' or 'USERF'), variant, scaling, padedegree, phiorder and an
optional function pointer; option names are matched case-insensitively, values
case-sensitively (serial_mod:97-101, 214-216).  `apply` evaluates
y = f(scaling*H) x for a small H (dense array, or a sparse matrix densified
first as psfun_d_serial_apply_sparse does) on the GPU:

  EXP    TAYLOR, SASQ, GENPADE, CHBHES, CHBGEN, CHBSYM (serial_mod:216-390)
  USERF  SYM: eigen-decomposition of the symmetric matrix of H's upper
         triangle on the device, f_ptr(scaling*lambda) on the host, as
         DSYEV('V','U') + f_ptr + DGEMV there (serial_mod:420-485)

'PHI' needs the external phi-function library, which the reference only
uses when built WITHPHILIBRARY; without it the reference reports an error,
and so does this apply.
"""
import contextlib
import ctypes as C

import numpy as np

from . import _lib
from . import psb

# Variants of fname 'EXP' that the reference accepts (serial_mod:216-390).
EXP_VARIANTS = ("TAYLOR", "SASQ", "GENPADE", "CHBHES", "CHBGEN", "CHBSYM")
# Degrees of the GENPADE approximant the device evaluates (psfun_set_pade_degree).
PADE_MAX = 30
# Variants of fname 'USERF' (serial_mod:420-498).
USERF_VARIANTS = ("SYM",)


class psfun_d_serial:
    def __init__(self):
        self.fname = "EXP"
        self.variant = "EXPOKIT"   # the reference's default, which its apply rejects (serial_mod:45, 387-390)
        self.scaling = 1.0
        self.padedegree = 6
        self.phiorder = 1
        self.f_ptr = None

    def set(self, what, val):
        """Returns info (psb_success_ or psb_err_invalid_args_combination_)."""
        key = str(what).upper()
        if isinstance(val, str):
            if key == "FNAME":
                self.fname = val
            elif key == "VARIANT":
                self.variant = val
            else:
                return psb.psb_err_invalid_args_combination_
        elif callable(val):
            if key == "FPOINTER":
                self.f_ptr = val
            else:
                return psb.psb_err_invalid_args_combination_
        elif isinstance(val, (int, np.integer)) and not isinstance(val, bool) and key in ("PADE_DEGREE", "PHIORDER"):
            if key == "PADE_DEGREE":
                self.padedegree = int(val)
            else:
                self.phiorder = int(val)
        elif isinstance(val, (float, int, np.floating, np.integer)):
            if key == "SCALING":
                self.scaling = float(val)
            else:
                return psb.psb_err_invalid_args_combination_
        else:
            return psb.psb_err_invalid_args_combination_
        return psb.psb_success_

    def device_fn(self):
        """(status, fname id, variant id) of the device f(H) step for this
        object (the select cases of serial_mod:214-505)."""
        if self.fname == "EXP":
            if self.variant not in EXP_VARIANTS:
                return _lib.PSFUN_ERR_INVALID_ARG, None, None
            if self.variant == "GENPADE" and not 2 <= self.padedegree <= PADE_MAX:
                return _lib.PSFUN_ERR_INVALID_ARG, None, None
            return _lib.PSFUN_SUCCESS, _lib.PSFUN_FN_EXP, _lib.VARIANTS[self.variant]
        if self.fname == "USERF":
            if self.variant not in USERF_VARIANTS or self.f_ptr is None:
                return _lib.PSFUN_ERR_INVALID_ARG, None, None
            return _lib.PSFUN_SUCCESS, _lib.PSFUN_FN_USERF, _lib.PSFUN_VAR_SYM
        return _lib.PSFUN_ERR_UNSUPPORTED, None, None

    @contextlib.contextmanager
    def bound(self, ctx):
        """Context manager that makes this object's f_ptr the context's USERF
        function, and its Pade degree the context's GENPADE degree
        (psfun_set_pade_degree; DGPADM takes fun%padedegree,
        serial_mod:271), for the calls inside it."""
        with contextlib.ExitStack() as stack:
            if self.fname == "USERF" and self.f_ptr is not None:
                stack.enter_context(ctx.userf(self.f_ptr))
            if self.fname == "EXP" and self.variant == "GENPADE" and self.padedegree != 6:
                _lib.check(_lib.lib().psfun_set_pade_degree(ctx.handle, int(self.padedegree)),
                           "psfun_set_pade_degree")
                stack.callback(_lib.lib().psfun_set_pade_degree, ctx.handle, 6)
            yield

    def apply(self, a, y, x, ctxt=None):
        """y[:] = f(scaling*a) x for a k x k matrix `a`: a dense host array
        (psfun_d_serial_apply_array) or a psb_dspmat_type / (rowptr, colind,
        val) CSR triple, densified first (psfun_d_serial_apply_sparse,
        serial_mod:526-573).  Returns info."""
        if isinstance(a, psb.psb_dspmat_type):
            a = a.host_csr()
        if isinstance(a, tuple):
            rp, ci, va = (np.asarray(t) for t in a)
            nrow = len(rp) - 1
            dense = np.zeros((nrow, nrow))
            for i in range(nrow):
                for q in range(int(rp[i]), int(rp[i + 1])):
                    dense[i, int(ci[q])] += va[q]   # repeated entries summed, as the SpMV sees them
            a = dense
        a = np.asarray(a, dtype=np.float64)
        if a.ndim != 2 or a.shape[0] != a.shape[1]:
            return psb.psb_err_from_subroutine_
        st, fn, var = self.device_fn()
        if st != _lib.PSFUN_SUCCESS:
            return psb.psb_err_from_subroutine_
        k = a.shape[0]
        H = np.ascontiguousarray(a.reshape(-1, order="F"))
        xx = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(k)
        ctx = ctxt if ctxt is not None else psb.default_context()
        with self.bound(ctx):
            st = _lib.lib().psfun_serial_apply_host(ctx.handle, fn, var, float(self.scaling), k,
                                                    H.ctypes.data_as(C.c_void_p), xx.ctypes.data_as(C.c_void_p),
                                                    out.ctypes.data_as(C.c_void_p))
        if st != _lib.PSFUN_SUCCESS:
            return psb.psb_err_from_subroutine_
        y[:] = out
        return psb.psb_success_
