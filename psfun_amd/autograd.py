"""Differentiable f(A)X: a torch.autograd.Function for Y = exp(-tA) X with
hand-written adjoint kernels (no reference counterpart; the reference is not
differentiable).

Forward:   Y = exp(-tA) X by the batched device Lanczos (one matrix, K columns).
Backward (A symmetric, loss L with dL/dY = G):
  dL/dX    = exp(-tA^T) G = exp(-tA) G          -- a second batched Lanczos
  dL/dt    = -<G, A Y>                          -- one SpMM + a dot
  dL/da_ik = -Re sum_j c_j sum_col u_j[i] v_j[k] -- sampled (SDDMM) over A's pattern
             with exp(-tA) ~ Re sum_j c_j (xi_j I + A)^{-1} (modified Talbot rule),
             v_j = (xi_j I + A)^{-1} x_col and u_j = (xi_j I + A)^{-T} g_col from the
             multi-shift CG (all poles of one column in one Krylov space).
The Krylov basis of the forward is not kept across the call: the backward
recomputes what it needs (a Lanczos run on G), so memory stays at one basis.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from . import psb
from . import quadrature as Q
from .batched import psfun_batch


class ExpmvOperator:
    """exp(-tA) on a fixed symmetric sparsity pattern (0-based CSR)."""

    def __init__(self, rowptr, colind, eps=1e-10, itmax=250, quad_poles=24, quad_eps=1e-13, quad_itmax=20000,
                 ctxt=None):
        self.ctxt = ctxt if ctxt is not None else psb.default_context()
        self.rowptr = np.ascontiguousarray(rowptr, np.int32)
        self.colind = np.ascontiguousarray(colind, np.int32)
        self.n = len(self.rowptr) - 1
        self.nnz = len(self.colind)
        ones = np.ones(self.nnz)
        self.batch = psfun_batch([(self.rowptr, self.colind, ones)], self.ctxt)
        self.mat = psb.psb_dspmat_type(self.rowptr, self.colind, ones, 0, self.ctxt)
        self.eps, self.itmax = eps, itmax
        self.quad_poles, self.quad_eps, self.quad_itmax = quad_poles, quad_eps, quad_itmax
        rp = C.c_void_p()
        ci = C.c_void_p()
        _lib.check(_lib.lib().psfun_batch_csr(self.batch.handle, C.byref(rp), C.byref(ci), None))
        self._d_rowptr, self._d_colind = rp, ci
        self.last_iters = None

    def set_values(self, vals):
        v = vals.detach().contiguous()
        assert v.is_cuda and v.dtype == torch.float64 and v.numel() == self.nnz
        self.ctxt.bind_stream()
        _lib.check(_lib.lib().psfun_batch_set_values(self.batch.handle, C.c_void_p(v.data_ptr())))
        _lib.check(_lib.lib().psfun_csr_set_values(self.mat.handle, C.c_void_p(v.data_ptr())))
        torch.cuda.current_stream().synchronize()

    def expm(self, X, t):
        X = X.contiguous()
        Y = torch.empty_like(X)
        iters, _ = self.batch.lanczos(X, Y, fscale=1.0, opscale=-float(t), eps=self.eps, itmax=self.itmax)
        self.last_iters = iters
        return Y

    def spmm(self, X):
        X = X.contiguous()
        Y = torch.empty_like(X)
        K = 1 if X.dim() == 1 else int(X.shape[1])
        self.ctxt.bind_stream()
        _lib.check(_lib.lib().psfun_batch_spmm(self.ctxt.handle, self.batch.handle, K, 1.0, C.c_void_p(X.data_ptr()),
                                               0.0, C.c_void_p(Y.data_ptr())))
        return Y

    def grad_values(self, X, G, t):
        """dL/da for every stored nonzero, L = <G, exp(-tA) X>."""
        xi, c, eta, sign, info = Q.talbot(None, self.quad_poles, rparams=[float(t)])
        shifts, weights = Q.fold_conjugate_pairs(sign * xi, c)
        S = len(shifts)
        sh = np.ascontiguousarray(np.stack([shifts.real, shifts.imag], 1).reshape(-1))
        ww = np.ascontiguousarray(np.stack([weights.real, weights.imag], 1).reshape(-1))
        grad = torch.zeros(self.nnz, dtype=torch.float64, device=X.device)
        K = 1 if X.dim() == 1 else int(X.shape[1])
        Xc = X.reshape(self.n, K)
        Gc = G.reshape(self.n, K)
        V = torch.empty((S, self.n), dtype=torch.complex128, device=X.device)
        U = torch.empty_like(V)
        L = _lib.lib()
        self.ctxt.bind_stream()
        it = np.zeros(S, dtype=np.int32)
        for col in range(K):
            for src, dst in ((Xc[:, col].contiguous(), V), (Gc[:, col].contiguous(), U)):
                st = L.psfun_mscg(self.ctxt.handle, self.mat.handle, S, sh.ctypes.data_as(C.c_void_p),
                                  C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()), self.n, float(self.quad_eps),
                                  int(self.quad_itmax), it.ctypes.data_as(C.c_void_p), None)
                _lib.check(st, "psfun_mscg")
            st = L.psfun_sddmm_shifted(self.ctxt.handle, self.n, self._d_rowptr, self._d_colind, S,
                                       ww.ctypes.data_as(C.c_void_p), C.c_void_p(U.data_ptr()), self.n,
                                       C.c_void_p(V.data_ptr()), self.n, -1.0, C.c_void_p(grad.data_ptr()))
            _lib.check(st, "psfun_sddmm_shifted")
        return grad


class ExpmvFunction(torch.autograd.Function):
    """Y = exp(-t A(vals)) X.  Inputs: vals (nnz,), X (n, K), t (0-dim tensor)."""

    @staticmethod
    def forward(ctx, vals, X, t, op):
        op.set_values(vals)
        Y = op.expm(X, float(t))
        ctx.op = op
        ctx.save_for_backward(vals, X, t, Y)
        return Y

    @staticmethod
    def backward(ctx, G):
        vals, X, t, Y = ctx.saved_tensors
        op = ctx.op
        G = G.contiguous()
        op.set_values(vals)
        tf = float(t)
        gvals = gX = gt = None
        if ctx.needs_input_grad[1]:
            gX = op.expm(G, tf)
        if ctx.needs_input_grad[2]:
            gt = -(G * op.spmm(Y)).sum().reshape(t.shape)
        if ctx.needs_input_grad[0]:
            gvals = op.grad_values(X, G, tf)
        return gvals, gX, gt, None


def expmv(vals, X, t, op):
    """Differentiable exp(-t A) X on the pattern held by `op`."""
    return ExpmvFunction.apply(vals, X, t, op)
