"""Minimal single-process PSBLAS surface used by PSFUN's f(A)x path, backed
by the gfx950 library.

Mirrors the PSBLAS objects and primitives that PSFUN's algorithms call
(krylov/polynomial/psfun_d_lanczos.f90, krylov/shifted/psfun_dcg.F90):
psb_desc_type, psb_dspmat_type (CSR), psb_d_vect_type, psb_spmm, psb_gedot,
psb_genrm2, psb_geamax, psb_geaxpby.  Vectors live on the GPU as float64
torch tensors; every primitive runs a HIP kernel of libpsfun_hip.so.

Error behaviour follows PSBLAS: routines return an integer `info`
(psb_success_ == 0) instead of raising.
"""
import contextlib
import ctypes as C

import numpy as np
import torch

from . import _lib

# PSBLAS error names used by the reference.  The numeric values are this
# package's own (PSBLAS is not part of the reference tree); callers compare
# against the names.
psb_success_ = 0
psb_err_invalid_istop_ = 5001
psb_err_invalid_vect_state_ = 1124
psb_err_invalid_args_combination_ = 1126
psb_err_from_subroutine_ = 4010
psb_err_from_subroutine_non_ = 4011
psb_err_alloc_dealloc_ = 4000
psb_err_unsupported_ = 4050

_STATUS_TO_INFO = {
    _lib.PSFUN_SUCCESS: psb_success_,
    _lib.PSFUN_ERR_INVALID_ISTOP: psb_err_invalid_istop_,
    _lib.PSFUN_ERR_INVALID_VECT_STATE: psb_err_invalid_vect_state_,
    _lib.PSFUN_ERR_INVALID_ARGS_COMBINATION: psb_err_invalid_args_combination_,
    _lib.PSFUN_ERR_ALLOC: psb_err_alloc_dealloc_,
    _lib.PSFUN_ERR_UNSUPPORTED: psb_err_unsupported_,
}


def status_to_info(status):
    return _STATUS_TO_INFO.get(status, psb_err_from_subroutine_non_)


class Context:
    """One libpsfun_hip context (stream-ordered with torch's current stream)."""

    def __init__(self, device=0):
        L = _lib.lib()
        if not torch.cuda.is_available():
            raise RuntimeError("psfun_amd needs a HIP GPU (torch.cuda.is_available() is False)")
        self.device = int(device)
        h = C.c_void_p()
        _lib.check(L.psfun_ctx_create(self.device, C.byref(h)), "psfun_ctx_create")
        self.handle = h

    def bind_stream(self):
        s = torch.cuda.current_stream(self.device).cuda_stream
        _lib.lib().psfun_ctx_set_stream(self.handle, C.c_void_p(s))

    @contextlib.contextmanager
    def userf(self, f):
        """Make f the USERF function (psfun_set_userf) for the calls inside
        the block.  The library calls it on this thread; an exception raised
        by f makes the remaining calls return NaN and is re-raised here."""
        errors = []

        def cb(x, _user):
            if errors:
                return float("nan")
            try:
                return float(f(x))
            except BaseException as e:  # noqa: BLE001  (re-raised after the call)
                errors.append(e)
                return float("nan")

        cfun = _lib.USERF(cb)
        _lib.check(_lib.lib().psfun_set_userf(self.handle, C.cast(cfun, C.c_void_p), None), "psfun_set_userf")
        try:
            yield
        finally:
            _lib.lib().psfun_set_userf(self.handle, None, None)
        if errors:
            raise errors[0]

    def num_cus(self):
        v = C.c_int(0)
        _lib.check(_lib.lib().psfun_ctx_num_cus(self.handle, C.byref(v)))
        return v.value

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib().psfun_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_contexts = {}


def log_trace(methdname, itrace, it, err, res, eps, estimates=True, out=None):
    """The convergence log of the reference's Krylov loops (log_header,
    log_conv every itrace-th estimate, log_end; psfun_d_lanczos.f90:180,
    304-305, 365), printed after the device run from its history: the
    estimate at loop index i is rs(i) = res[i-1], logged for itx = i - 1
    (itx = it is the final err).  Without estimates (istop 2) only the
    header and the end line appear, as log_conv is not called then.
    PSBLAS's own line format is not part of the reference sources; this one
    is the Fortran layer's."""
    if itrace is None or int(itrace) <= 0:
        return
    import sys
    w = (out or sys.stdout).write
    tr = int(itrace)
    w(f"{methdname}:   Iteration   Error estimate   Tolerance\n")
    if estimates:
        for i in range(2, it + 2):
            if (i - 1) % tr == 0 and i - 1 < len(res):
                w(f"{methdname}:{i - 1:6d}  {res[i - 1]:15.8E}  {eps:15.8E}\n")
    w(f"{methdname}: iterations {it:6d}, error estimate {err:15.8E}\n")


def default_context(device=0):
    if device not in _contexts:
        _contexts[device] = Context(device)
    ctx = _contexts[device]
    ctx.bind_stream()
    return ctx


class psb_desc_type:
    """Descriptor of a single-process row distribution of n rows."""

    def __init__(self, n, ctxt=None):
        self.n = int(n)
        self.ctxt = ctxt if ctxt is not None else default_context()

    def get_context(self):
        return self.ctxt

    def get_global_rows(self):
        return self.n

    def get_local_rows(self):
        return self.n

    def get_local_cols(self):
        return self.n


def _dptr(t):
    return C.c_void_p(t.data_ptr())


class psb_dspmat_type:
    """Square CSR matrix resident in HBM (int32 indices, FP64 values).

    Built from host arrays (rowptr, colind, val), 0- or 1-based (PSBLAS CSR
    is 1-based); they are uploaded once and converted to 0-based."""

    def __init__(self, rowptr, colind, val, index_base=0, ctxt=None):
        self.ctxt = ctxt if ctxt is not None else default_context()
        rp = np.ascontiguousarray(rowptr, dtype=np.int32)
        ci = np.ascontiguousarray(colind, dtype=np.int32)
        va = np.ascontiguousarray(val, dtype=np.float64)
        self.n = len(rp) - 1
        self.nnz = len(ci)
        h = C.c_void_p()
        st = _lib.lib().psfun_csr_create(self.ctxt.handle, self.n, self.nnz, rp.ctypes.data_as(C.c_void_p),
                                          ci.ctypes.data_as(C.c_void_p), va.ctypes.data_as(C.c_void_p),
                                          int(index_base), _lib.PSFUN_MEM_HOST, C.byref(h))
        _lib.check(st, "psfun_csr_create")
        self.handle = h
        self._src = (rp, ci, va, int(index_base))  # the host arrays as given, for host_csr()

    def host_csr(self):
        """0-based host (rowptr, colind, val) of this matrix, with the values
        last given (at creation or by set_values)."""
        rp, ci, va, base = self._src
        if isinstance(va, torch.Tensor):
            va = va.detach().cpu().numpy()
        return rp - base, ci - base, np.asarray(va, dtype=np.float64)

    @classmethod
    def from_csr(cls, csr, ctxt=None):
        rp, ci, va = csr
        return cls(rp, ci, va, 0, ctxt)

    def get_nrows(self):
        return self.n

    def get_ncols(self):
        return self.n

    def get_nzeros(self):
        return self.nnz

    def set_lanes(self, lanes):
        _lib.check(_lib.lib().psfun_csr_set_lanes(self.handle, int(lanes)))

    def lanes(self):
        v = C.c_int(0)
        _lib.check(_lib.lib().psfun_csr_info(self.handle, None, None, C.byref(v)))
        return v.value

    def set_values(self, vals):
        """Replace the nonzero values (same pattern) from a float64 device tensor."""
        v = vals.detach().contiguous()
        if not (v.is_cuda and v.dtype == torch.float64 and v.numel() == self.nnz):
            raise ValueError("set_values needs a float64 device tensor with nnz entries")
        self.ctxt.bind_stream()
        _lib.check(_lib.lib().psfun_csr_set_values(self.handle, _dptr(v)))
        torch.cuda.current_stream().synchronize()
        self._src = self._src[:2] + (v, self._src[3])

    _FORMATS = {"AUTO": _lib.PSFUN_FORMAT_AUTO, "CSR": _lib.PSFUN_FORMAT_CSR, "SELL": _lib.PSFUN_FORMAT_SELL,
                "HDIA": _lib.PSFUN_FORMAT_HDIA, "CSRS": _lib.PSFUN_FORMAT_CSRS}

    def set_format(self, fmt):
        """'CSR' (vector CSR), 'SELL' (sliced ELL, 64-row slices), 'HDIA'
        (diagonal slices, as PSBLAS's hacked DIA; raises PsfunError with
        PSFUN_ERR_UNSUPPORTED when the slices would hold more than 1.5 x nnz
        values) or 'AUTO'."""
        self.ctxt.bind_stream()
        _lib.check(_lib.lib().psfun_csr_set_format(self.handle, self._FORMATS[str(fmt).upper()]))

    def hdia_diagonals(self):
        """Slice diagonals of the diagonal-slice copy (0 when not built)."""
        d = C.c_int64(0)
        _lib.check(_lib.lib().psfun_csr_hdia_diagonals(self.handle, C.byref(d)))
        return d.value

    def sell_index_bits(self):
        """16 or 32: width of the column indices the Lanczos SELL step reads
        (0 without a sliced-ELL copy)."""
        b = C.c_int(0)
        _lib.check(_lib.lib().psfun_csr_sell_index_bits(self.handle, C.byref(b)))
        return b.value

    def get_format(self):
        """(format name, padded sliced-ELL entry count)."""
        f = C.c_int(0)
        e = C.c_int64(0)
        _lib.check(_lib.lib().psfun_csr_get_format(self.handle, C.byref(f), C.byref(e)))
        name = {v: k for k, v in self._FORMATS.items()}[f.value]
        return name, e.value

    def free(self):
        if getattr(self, "handle", None):
            _lib.lib().psfun_csr_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class psb_d_vect_type:
    """Distributed (here: single-GPU) FP64 vector.  `v` is None until allocated."""

    def __init__(self, v=None):
        self.v = v

    @classmethod
    def from_numpy(cls, a, device=0):
        return cls(torch.as_tensor(np.ascontiguousarray(a, np.float64), device=f"cuda:{device}"))

    @classmethod
    def zeros(cls, n, device=0):
        return cls(torch.zeros(int(n), dtype=torch.float64, device=f"cuda:{device}"))

    def get_nrows(self):
        return 0 if self.v is None else int(self.v.numel())

    def get_vect(self):
        return self.v.detach().cpu().numpy().copy()

    def is_allocated(self):
        return self.v is not None


def _vec_ok(*vs):
    for v in vs:
        if v is None or v.v is None:
            return False
        t = v.v
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
            return False
    return True


def psb_spmm(alpha, a, x, beta, y, desc_a=None):
    """y <- alpha*A*x + beta*y; returns info."""
    if not _vec_ok(x, y):
        return psb_err_invalid_vect_state_
    ctx = a.ctxt
    ctx.bind_stream()
    return status_to_info(_lib.lib().psfun_spmv(ctx.handle, a.handle, float(alpha), _dptr(x.v), float(beta),
                                                 _dptr(y.v)))


def psb_gedot(x, y, desc_a):
    """(x . y, info)"""
    if not _vec_ok(x, y):
        return 0.0, psb_err_invalid_vect_state_
    ctx = desc_a.get_context()
    ctx.bind_stream()
    r = C.c_double(0.0)
    st = _lib.lib().psfun_dot(ctx.handle, x.get_nrows(), _dptr(x.v), _dptr(y.v), C.byref(r))
    return r.value, status_to_info(st)


def psb_genrm2(x, desc_a):
    if not _vec_ok(x):
        return 0.0, psb_err_invalid_vect_state_
    ctx = desc_a.get_context()
    ctx.bind_stream()
    r = C.c_double(0.0)
    st = _lib.lib().psfun_nrm2(ctx.handle, x.get_nrows(), _dptr(x.v), C.byref(r))
    return r.value, status_to_info(st)


def psb_geamax(x, desc_a):
    if not _vec_ok(x):
        return 0.0, psb_err_invalid_vect_state_
    ctx = desc_a.get_context()
    ctx.bind_stream()
    r = C.c_double(0.0)
    st = _lib.lib().psfun_amax(ctx.handle, x.get_nrows(), _dptr(x.v), C.byref(r))
    return r.value, status_to_info(st)


def psb_geaxpby(alpha, x, beta, y, desc_a):
    """y <- alpha*x + beta*y; returns info."""
    if not _vec_ok(x, y):
        return psb_err_invalid_vect_state_
    ctx = desc_a.get_context()
    ctx.bind_stream()
    return status_to_info(_lib.lib().psfun_axpby(ctx.handle, x.get_nrows(), float(alpha), _dptr(x.v),
                                                  float(beta), _dptr(y.v)))


def psb_scal(alpha, x, desc_a):
    """x <- alpha*x (v%scal); returns info."""
    if not _vec_ok(x):
        return psb_err_invalid_vect_state_
    ctx = desc_a.get_context()
    ctx.bind_stream()
    return status_to_info(_lib.lib().psfun_scal(ctx.handle, x.get_nrows(), float(alpha), _dptr(x.v)))
