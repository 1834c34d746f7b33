"""psfun_d_serial: the function object of the f(H) step
(serial/psfun_d_serial_mod.F90:43-170).

Holds fname ('EXP'), variant, scaling, padedegree, phiorder and an optional
function pointer; option names are matched case-insensitively, values
case-sensitively (serial_mod:97-101, 214-216).  `apply` evaluates
y = f(scaling*H) x for a small dense H on the GPU.
"""
import ctypes as C

import numpy as np

from . import _lib
from . import psb

# Variants of fname 'EXP' that the reference accepts (serial_mod:216-390).
EXP_VARIANTS = ("TAYLOR", "SASQ", "GENPADE", "CHBHES", "CHBGEN", "CHBSYM")


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

    def device_variant(self):
        """(status, variant id) of the device f(H) step for this object."""
        if self.fname != "EXP":
            return _lib.PSFUN_ERR_UNSUPPORTED, None
        if self.variant not in EXP_VARIANTS:
            return _lib.PSFUN_ERR_INVALID_ARG, None
        return _lib.PSFUN_SUCCESS, _lib.VARIANTS[self.variant]

    def apply(self, a, y, x, ctxt=None):
        """y[:] = f(scaling*a) x for a dense k x k array `a` (host numpy).
        Returns info."""
        a = np.asarray(a, dtype=np.float64)
        if a.ndim != 2 or a.shape[0] != a.shape[1]:
            return psb.psb_err_from_subroutine_
        st, var = self.device_variant()
        if st != _lib.PSFUN_SUCCESS:
            return psb.psb_err_from_subroutine_
        k = a.shape[0]
        H = np.ascontiguousarray(a.reshape(-1, order="F"))
        xx = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(k)
        ctx = ctxt if ctxt is not None else psb.default_context()
        st = _lib.lib().psfun_serial_apply_host(ctx.handle, _lib.PSFUN_FN_EXP, var, float(self.scaling), k,
                                                H.ctypes.data_as(C.c_void_p), xx.ctypes.data_as(C.c_void_p),
                                                out.ctypes.data_as(C.c_void_p))
        if st != _lib.PSFUN_SUCCESS:
            return psb.psb_err_from_subroutine_
        y[:] = out
        return psb.psb_success_
