"""Batched f(A)x: many independent matrices of uneven order, each with a
block of K vectors, advanced together on the GPU (one launch sequence,
optionally captured once in a hipGraph and replayed).

  B = psfun_batch([csr_0, csr_1, ...])        # csr_g = (rowptr, colind, val), 0-based
  Y, iters, errs = B.expm(X, t=0.5, eps=1e-10, itmax=50)   # X: (sum n_g) x K row-major

Every (graph, column) pair is one run of the reference's Lanczos
(krylov/polynomial/psfun_d_lanczos.f90) with its own tridiagonal matrix,
stopping test and assembly.  The launch shape (rows per block-chunk, hipGraph
on/off) is chosen by `autotune`, which times the candidates on the actual
batch and remembers the winner per batch shape.
"""
import ctypes as C
import time

import numpy as np
import torch

from . import _lib
from . import psb


def concat_block_diagonal(mats):
    """Concatenate 0-based CSR matrices into one block-diagonal CSR with
    global column indices; returns (row_offsets, rowptr, colind, val)."""
    offs = [0]
    rps, cis, vas = [], [], []
    nnz = 0
    for rp, ci, va in mats:
        n = len(rp) - 1
        rps.append(np.asarray(rp[:-1], np.int64) + nnz)
        cis.append(np.asarray(ci, np.int64) + offs[-1])
        vas.append(np.asarray(va, np.float64))
        nnz += int(rp[-1])
        offs.append(offs[-1] + n)
    rowptr = np.concatenate(rps + [np.array([nnz], np.int64)]).astype(np.int32)
    return (np.array(offs, np.int64), rowptr, np.concatenate(cis).astype(np.int32), np.concatenate(vas))


_TUNE_CACHE = {}


_ORDERS = {"natural": _lib.PSFUN_ORDER_NATURAL, "rcm": _lib.PSFUN_ORDER_RCM}


def rcm_order(rowptr, colind, base=0):
    """Reverse Cuthill-McKee order of one graph (psfun_rcm_order, host):
    perm[p] = original row of new row p."""
    rp = np.ascontiguousarray(rowptr, np.int32)
    ci = np.ascontiguousarray(colind, np.int32)
    n = len(rp) - 1
    perm = np.zeros(max(n, 1), np.int32)
    _lib.check(_lib.lib().psfun_rcm_order(n, rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p), int(base),
                                          perm.ctypes.data_as(C.c_void_p)), "psfun_rcm_order")
    return perm[:n]


class psfun_batch:
    def __init__(self, mats, ctxt=None, order="natural"):
        """order 'natural' keeps every graph's row numbering on the device;
        'rcm' renumbers each graph by reverse Cuthill-McKee for gather
        locality (X, Y and values stay in the caller's numbering)."""
        self.ctxt = ctxt if ctxt is not None else psb.default_context()
        self.row_offsets, rp, ci, va = concat_block_diagonal(mats)
        self.ngraph = len(mats)
        self.nrows = int(self.row_offsets[-1])
        self.nnz = int(rp[-1])
        self.order = str(order).lower()
        h = C.c_void_p()
        st = _lib.lib().psfun_batch_create_ordered(self.ctxt.handle, self.ngraph,
                                                   self.row_offsets.ctypes.data_as(C.c_void_p),
                                                   rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p),
                                                   va.ctypes.data_as(C.c_void_p), _ORDERS[self.order], C.byref(h))
        _lib.check(st, "psfun_batch_create_ordered")
        self.handle = h
        self.rows_per_chunk = 512
        self.use_graph = False

    def set_tuning(self, rows_per_chunk, use_graph):
        _lib.check(_lib.lib().psfun_batch_set_tuning(self.handle, int(rows_per_chunk), int(bool(use_graph))))
        self.rows_per_chunk = int(rows_per_chunk)
        self.use_graph = bool(use_graph)

    def lanczos(self, X, Y, fscale=1.0, opscale=1.0, eps=1e-10, itmax=50, istop=1, want_stats=True):
        """Y = exp(fscale*opscale*A_g) X per graph and column (device tensors,
        row-major (nrows, K) float64).  Returns (iters, errs) or (None, None)."""
        assert X.is_cuda and X.dtype == torch.float64 and X.is_contiguous() and X.shape[0] == self.nrows
        K = 1 if X.dim() == 1 else int(X.shape[1])
        ninst = self.ngraph * K
        iters = np.zeros(ninst, dtype=np.int32) if want_stats else None
        errs = np.zeros(ninst) if want_stats else None
        self.ctxt.bind_stream()
        st = _lib.lib().psfun_batch_lanczos(self.ctxt.handle, self.handle, K, float(fscale), float(opscale),
                                            C.c_void_p(X.data_ptr()), C.c_void_p(Y.data_ptr()), float(eps), int(itmax),
                                            int(istop), iters.ctypes.data_as(C.c_void_p) if want_stats else None,
                                            errs.ctypes.data_as(C.c_void_p) if want_stats else None)
        _lib.check(st, "psfun_batch_lanczos")
        if want_stats:
            return iters.reshape(self.ngraph, K), errs.reshape(self.ngraph, K)
        return None, None

    def expm(self, X, t=1.0, eps=1e-10, itmax=50):
        """exp(-t A_g) X for every graph (the Laplacian heat kernel)."""
        Y = torch.empty_like(X)
        iters, errs = self.lanczos(X, Y, fscale=1.0, opscale=-float(t), eps=eps, itmax=itmax)
        return Y, iters, errs

    def autotune(self, X, t=1.0, eps=1e-10, itmax=50, candidates=(128, 256, 512, 1024, 2048), reps=2):
        """Time every (rows_per_chunk, hipGraph) candidate on this batch and
        keep the fastest.  Results are cached per (nrows, nnz, ngraph, K)."""
        K = 1 if X.dim() == 1 else int(X.shape[1])
        key = (self.nrows, self.nnz, self.ngraph, K, itmax, self.order)
        if key in _TUNE_CACHE:
            self.set_tuning(*_TUNE_CACHE[key][0])
            return _TUNE_CACHE[key]
        Y = torch.empty_like(X)
        best = None
        table = []
        for rpc in candidates:
            for g in (False, True):
                self.set_tuning(rpc, g)
                self.lanczos(X, Y, opscale=-float(t), eps=eps, itmax=itmax, want_stats=False)  # warm (and capture)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(reps):
                    self.lanczos(X, Y, opscale=-float(t), eps=eps, itmax=itmax, want_stats=False)
                torch.cuda.synchronize()
                dt = (time.perf_counter() - t0) / reps
                table.append(((rpc, g), dt))
                if best is None or dt < best[1]:
                    best = ((rpc, g), dt)
        self.set_tuning(*best[0])
        _TUNE_CACHE[key] = (best[0], table)
        return _TUNE_CACHE[key]

    def free(self):
        if getattr(self, "handle", None):
            _lib.lib().psfun_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
