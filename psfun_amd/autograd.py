"""Differentiable f(A)X: a torch.autograd.Function for Y = exp(-tA) X with
hand-written adjoint kernels (no reference counterpart; the reference is not
differentiable).

Forward:  Y = exp(-tA) X by the batched device Lanczos (one matrix, K columns),
          keeping the Krylov basis V of every column (or not: `store_basis`).
Backward (A symmetric, loss L with dL/dY = G):
  dL/dX    = exp(-tA^T) G = exp(-tA) G        -- a second batched Lanczos, on G,
                                                 which also yields its basis W
  dL/dt    = -<G, A Y>                        -- one SpMM + a dot
  dL/da_ik = -Re sum_s w_s sum_col u_s[i] v_s[k] -- SDDMM over A's pattern, with the
             resolvent form exp(-tA) ~ Re sum_s w_s (xi_s I + A)^{-1} (modified Talbot
             rule), v_s = (xi_s I + A)^{-1} x_col and u_s = (xi_s I + A)^{-T} g_col taken
             from the bases: v_s = ||x|| V_m (xi_s I + T_m)^{-1} e1 (the multi-shift CG
             iterate of step m), likewise u_s from W.
With store_basis=False the forward keeps nothing and the backward rebuilds V
by re-running the forward Lanczos (memory for one basis at a time);
store_basis="auto" (the default) keeps the basis only when the forward and
adjoint bases fit in half of the device memory that is free at forward time
(including what torch's caching allocator holds unused).
grad_mode="mscg" takes v_s, u_s from full multi-shift CG solves instead (one
column at a time; slower, used to cross-check).
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

    def __init__(self, rowptr, colind, eps=1e-10, itmax=150, quad_poles=24, quad_eps=1e-13, quad_itmax=20000,
                 store_basis="auto", grad_mode="krylov", ctxt=None):
        self.ctxt = ctxt if ctxt is not None else psb.default_context()
        self.rowptr = np.ascontiguousarray(rowptr, np.int32)
        self.colind = np.ascontiguousarray(colind, np.int32)
        self.n = len(self.rowptr) - 1
        self.nnz = len(self.colind)
        ones = np.ones(self.nnz)
        self.batch = psfun_batch([(self.rowptr, self.colind, ones)], self.ctxt)
        self.mat = None
        if grad_mode == "mscg":
            self.mat = psb.psb_dspmat_type(self.rowptr, self.colind, ones, 0, self.ctxt)
        self.eps, self.itmax = eps, itmax
        self.quad_poles, self.quad_eps, self.quad_itmax = quad_poles, quad_eps, quad_itmax
        self.store_basis = store_basis
        self.grad_mode = grad_mode
        rp = C.c_void_p()
        ci = C.c_void_p()
        _lib.check(_lib.lib().psfun_batch_csr(self.batch.handle, C.byref(rp), C.byref(ci), None))
        self._d_rowptr, self._d_colind = rp, ci
        self.last_iters = None

    # ---- primitives ------------------------------------------------------
    def set_values(self, vals):
        v = vals.detach().contiguous()
        assert v.is_cuda and v.dtype == torch.float64 and v.numel() == self.nnz
        self.ctxt.bind_stream()
        _lib.check(_lib.lib().psfun_batch_set_values(self.batch.handle, C.c_void_p(v.data_ptr())))
        if self.mat is not None:
            _lib.check(_lib.lib().psfun_csr_set_values(self.mat.handle, C.c_void_p(v.data_ptr())))
        torch.cuda.current_stream().synchronize()

    def basis_bytes(self, K):
        """Device bytes of one stored Krylov basis for K columns."""
        return 8 * (self.itmax - 1) * self.n * K + 16 * K * (self.itmax + 2)

    def keep_basis(self, K, device):
        """Whether the forward keeps its basis for the backward."""
        if self.store_basis != "auto":
            return bool(self.store_basis)
        free, _total = torch.cuda.mem_get_info(device)
        # memory torch's caching allocator holds but does not use is free for us too
        free += torch.cuda.memory_reserved(device) - torch.cuda.memory_allocated(device)
        # the backward holds the adjoint basis too, so two bases must fit
        return 2 * self.basis_bytes(K) <= 0.5 * free

    def expm(self, X, t):
        X = X.contiguous()
        Y = torch.empty_like(X)
        iters, _ = self.batch.lanczos(X, Y, fscale=1.0, opscale=-float(t), eps=self.eps, itmax=self.itmax)
        self.last_iters = iters
        return Y

    def expm_basis(self, X, t):
        """Y = exp(-tA) X and the Krylov data (V, T_alpha, T_beta, nbasis, rs1)."""
        X = X.contiguous()
        K = 1 if X.dim() == 1 else int(X.shape[1])
        dev = X.device
        Y = torch.empty_like(X)
        V = torch.empty(((self.itmax - 1) * self.n * K,), dtype=torch.float64, device=dev)
        ld = self.itmax + 2
        Ta = torch.empty((K * ld,), dtype=torch.float64, device=dev)
        Tb = torch.empty_like(Ta)
        nb = torch.empty((K,), dtype=torch.int32, device=dev)
        rs1 = torch.empty((K,), dtype=torch.float64, device=dev)
        self.ctxt.bind_stream()
        st = _lib.lib().psfun_batch_lanczos_basis(self.ctxt.handle, self.batch.handle, K, 1.0, -float(t),
                                                  C.c_void_p(X.data_ptr()), C.c_void_p(Y.data_ptr()), float(self.eps),
                                                  int(self.itmax), 1, C.c_void_p(V.data_ptr()),
                                                  C.c_void_p(Ta.data_ptr()), C.c_void_p(Tb.data_ptr()),
                                                  C.c_void_p(nb.data_ptr()), C.c_void_p(rs1.data_ptr()))
        _lib.check(st, "psfun_batch_lanczos_basis")
        self.last_iters = nb.cpu().numpy()
        return Y, (V, Ta, Tb, nb, rs1, K)

    def spmm(self, X):
        X = X.contiguous()
        Y = torch.empty_like(X)
        K = 1 if X.dim() == 1 else int(X.shape[1])
        self.ctxt.bind_stream()
        _lib.check(_lib.lib().psfun_batch_spmm(self.ctxt.handle, self.batch.handle, K, 1.0, C.c_void_p(X.data_ptr()),
                                               0.0, C.c_void_p(Y.data_ptr())))
        return Y

    def _rule(self, t):
        xi, c, eta, sign, info = Q.talbot(None, self.quad_poles, rparams=[float(t)])
        shifts, weights = Q.fold_conjugate_pairs(sign * xi, c)
        sh = np.ascontiguousarray(np.stack([shifts.real, shifts.imag], 1).reshape(-1))
        ww = np.ascontiguousarray(np.stack([weights.real, weights.imag], 1).reshape(-1))
        return len(shifts), sh, ww

    def resolvents(self, krylov, t):
        """Galerkin resolvents of every folded pole from stored Krylov data
        (the runs used the operator -t*A; their tridiagonals scale by -1/t)."""
        V, Ta, Tb, nb, rs1, K = krylov
        S, sh, _ = self._rule(t)
        mmax = int(nb.max().item())
        out = torch.empty((S, self.n, K), dtype=torch.complex128, device=V.device)
        self.ctxt.bind_stream()
        st = _lib.lib().psfun_krylov_resolvents(self.ctxt.handle, self.n, K, max(mmax, 1), C.c_void_p(V.data_ptr()),
                                                C.c_void_p(Ta.data_ptr()), C.c_void_p(Tb.data_ptr()), self.itmax + 2,
                                                -1.0 / float(t), C.c_void_p(nb.data_ptr()),
                                                C.c_void_p(rs1.data_ptr()), S,
                                                sh.ctypes.data_as(C.c_void_p), C.c_void_p(out.data_ptr()))
        _lib.check(st, "psfun_krylov_resolvents")
        return out

    def sddmm(self, U, Vr, t, grad, scale=-1.0):
        S, _, ww = self._rule(t)
        K = int(U.shape[2])
        self.ctxt.bind_stream()
        st = _lib.lib().psfun_sddmm_cols(self.ctxt.handle, self.n, self._d_rowptr, self._d_colind, K, S,
                                         ww.ctypes.data_as(C.c_void_p), C.c_void_p(U.data_ptr()),
                                         C.c_void_p(Vr.data_ptr()), float(scale), C.c_void_p(grad.data_ptr()))
        _lib.check(st, "psfun_sddmm_cols")
        return grad

    def vjp(self, X, Y, G, t, need_x=True, need_t=True, need_a=True):
        """(dL/dX, dL/dt, dL/dvals) for dL/dY = G in one C-ABI call
        (psfun_expv_vjp: the backward below with the forward basis
        recomputed); entries not asked for are None."""
        K = 1 if X.dim() == 1 else int(X.shape[1])
        X, Y, G = X.contiguous(), Y.contiguous(), G.contiguous()
        gX = torch.empty_like(G) if need_x else None
        gv = torch.empty(self.nnz, dtype=torch.float64, device=G.device) if need_a else None
        gt = C.c_double(0.0)
        self.ctxt.bind_stream()
        st = _lib.lib().psfun_expv_vjp(self.ctxt.handle, self.batch.handle, K, float(t), C.c_void_p(X.data_ptr()),
                                       C.c_void_p(Y.data_ptr()), C.c_void_p(G.data_ptr()), float(self.eps),
                                       int(self.itmax), int(self.quad_poles),
                                       C.c_void_p(gX.data_ptr()) if need_x else None,
                                       C.byref(gt) if need_t else None,
                                       C.c_void_p(gv.data_ptr()) if need_a else None)
        _lib.check(st, "psfun_expv_vjp")
        return gX, (gt.value if need_t else None), gv

    # ---- exact-solve variant (cross-check) --------------------------------
    def grad_values_mscg(self, X, G, t):
        S, sh, ww = self._rule(t)
        grad = torch.zeros(self.nnz, dtype=torch.float64, device=X.device)
        K = 1 if X.dim() == 1 else int(X.shape[1])
        Xc, Gc = X.reshape(self.n, K), G.reshape(self.n, K)
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
        tf = float(t)
        need_a = ctx.needs_input_grad[0] and op.grad_mode == "krylov"
        K = 1 if X.dim() == 1 else int(X.shape[1])
        if need_a and op.keep_basis(K, X.device):
            Y, kry = op.expm_basis(X, tf)
            ctx.krylov = kry
        else:
            Y = op.expm(X, tf)
            ctx.krylov = None
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
        need_a = ctx.needs_input_grad[0]
        if need_a and op.grad_mode == "krylov":
            gX, kry_g = op.expm_basis(G, tf)          # dL/dX and the adjoint basis W
            U = op.resolvents(kry_g, tf)
            del kry_g
            kry_x = ctx.krylov
            if kry_x is None:                          # rebuild the forward basis
                _, kry_x = op.expm_basis(X, tf)
            Vr = op.resolvents(kry_x, tf)
            del kry_x
            ctx.krylov = None
            gvals = torch.zeros(op.nnz, dtype=torch.float64, device=X.device)
            op.sddmm(U, Vr, tf, gvals, -1.0)
            del U, Vr
        else:
            if ctx.needs_input_grad[1]:
                gX = op.expm(G, tf)
            if need_a:
                gvals = op.grad_values_mscg(X, G, tf)
        if not ctx.needs_input_grad[1]:
            gX = None
        if ctx.needs_input_grad[2]:
            gt = -(G * op.spmm(Y)).sum().reshape(t.shape)
        return gvals, gX, gt, None


def expmv(vals, X, t, op):
    """Differentiable exp(-t A) X on the pattern held by `op`."""
    return ExpmvFunction.apply(vals, X, t, op)
