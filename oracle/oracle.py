"""ctypes front-end of the CPU oracle (oracle/psfun_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the CPU baseline,
never as the thing measured or shipped.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsfun_oracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lib = None

ZFUN = C.CFUNCTYPE(None, C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double))
DFUN = C.CFUNCTYPE(C.c_double, C.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from oracle import build as _b  # noqa: F401  (build on demand in the container)
            _b.build(verbose=False)
        L = C.CDLL(LIB_PATH)
        L.or_set_threads.argtypes = [C.c_int]
        L.or_get_max_threads.restype = C.c_int
        L.or_get_threads.restype = C.c_int
        L.or_spmv.argtypes = [C.c_int, _ip, _ip, _dp, C.c_double, _dp, C.c_double, _dp]
        L.or_dot.argtypes = [C.c_int, _dp, _dp]
        L.or_dot.restype = C.c_double
        L.or_nrm2.argtypes = [C.c_int, _dp]
        L.or_nrm2.restype = C.c_double
        L.or_amax.argtypes = [C.c_int, _dp]
        L.or_amax.restype = C.c_double
        L.or_axpby.argtypes = [C.c_int, C.c_double, _dp, C.c_double, _dp]
        L.or_tridiag_eig.argtypes = [C.c_int, _dp, _dp, _dp]
        L.or_tridiag_eig.restype = C.c_int
        L.or_exp_tridiag.argtypes = [C.c_int, _dp, _dp, C.c_double, _dp, _dp]
        L.or_exp_tridiag.restype = C.c_int
        L.or_dense_expm.argtypes = [C.c_int, _dp, C.c_double, _dp]
        L.or_dense_expm.restype = C.c_int
        kry = [C.c_int, _ip, _ip, _dp, C.c_double, C.c_double, _dp, _dp, C.c_double, C.c_int, C.c_int,
               C.POINTER(C.c_int), C.POINTER(C.c_double), _dp, C.POINTER(C.c_int)]
        L.or_lanczos.argtypes = kry
        L.or_lanczos.restype = C.c_int
        L.or_arnoldi.argtypes = kry
        L.or_arnoldi.restype = C.c_int
        _lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
        L.or_lanczos_batch.argtypes = [C.c_int, _lp, _ip, _ip, _dp, C.c_int, C.c_double, C.c_double, _dp, _dp,
                                       C.c_double, C.c_int, C.c_int, _ip, _dp]
        L.or_lanczos_batch.restype = C.c_int
        L.or_lanczos_f.argtypes = kry + [DFUN]
        L.or_lanczos_f.restype = C.c_int
        L.or_arnoldi_f.argtypes = kry + [DFUN]
        L.or_arnoldi_f.restype = C.c_int
        L.or_sym_eig.argtypes = [C.c_int, _dp, C.c_int, _dp, _dp]
        L.or_sym_eig.restype = C.c_int
        L.or_userf_sym_apply.argtypes = [C.c_int, _dp, C.c_int, C.c_double, DFUN, _dp, _dp]
        L.or_userf_sym_apply.restype = C.c_int
        L.or_sai_arnoldi.argtypes = [C.c_int, _ip, _ip, _dp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                     _dp, _dp, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int),
                                     C.POINTER(C.c_double), _dp, C.POINTER(C.c_int)]
        L.or_sai_arnoldi.restype = C.c_int
        L.or_sai_arnoldi_f.argtypes = list(L.or_sai_arnoldi.argtypes) + [DFUN]
        L.or_sai_arnoldi_f.restype = C.c_int
        L.or_dense_inverse.argtypes = [C.c_int, _dp, _dp]
        L.or_dense_inverse.restype = C.c_int
        L.or_cg_shifted.argtypes = [C.c_int, _ip, _ip, _dp, C.c_int, _dp, C.c_double, C.c_double, _dp, C.c_double,
                                    C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]
        L.or_cg_shifted.restype = C.c_int
        L.or_ilu0.argtypes = [C.c_int, _ip, _ip, _dp, _ip, _ip, _dp, C.POINTER(C.c_int)]
        L.or_ilu0.restype = C.c_int
        L.or_ilu0_apply.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _dp]
        L.or_ilu0_apply.restype = C.c_int
        L.or_mscg.argtypes = [C.c_int, _ip, _ip, _dp, C.c_int, _dp, _dp, _dp, C.c_double, C.c_int,
                              np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS"), _dp]
        L.or_mscg.restype = C.c_int
        L.or_ellipkkp.argtypes = [C.c_double, _dp]
        L.or_ellipj.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
        L.or_horner.argtypes = [C.c_int, _dp, C.c_double]
        L.or_horner.restype = C.c_double
        L.or_hhtmethod1.argtypes = [ZFUN, C.c_int, C.c_double, C.c_double, _dp, _dp, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]
        L.or_hhtmethod1.restype = C.c_int
        L.or_talbot.argtypes = [C.c_int, C.c_double, _dp, _dp]
        L.or_talbot.restype = C.c_int
        L.or_quad_apply.argtypes = [C.c_int, _ip, _ip, _dp, C.c_int, _dp, _dp, C.c_double, C.c_int, _dp, _dp,
                                    C.c_double, C.c_int, np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")]
        L.or_quad_apply.restype = C.c_int
        L.or_pade_expm.argtypes = [C.c_int, C.c_int, _dp, C.c_double, _dp]
        L.or_pade_expm.restype = C.c_int
        L.or_set_exp_cram.argtypes = [C.c_int]
        L.or_set_exp_cram.restype = None
        L.or_cram_apply.argtypes = [C.c_int, _dp, C.c_int, C.c_double, _dp]
        L.or_cram_apply.restype = C.c_int
        L.or_set_pade_degree.argtypes = [C.c_int]
        _i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        L.or_quad_apply_prec.argtypes = [C.c_int, _ip, _ip, _dp, C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_int,
                                         _dp, _dp, C.c_double, C.c_int, _i32, _dp]
        L.or_quad_apply_prec.restype = C.c_int
        L.or_pcocg.argtypes = [C.c_int, _ip, _ip, _dp, C.c_int, C.c_double, C.c_double, _dp, _dp, C.c_double,
                               C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.or_pcocg.restype = C.c_int
        L.or_mscg_prec.argtypes = [C.c_int, _ip, _ip, _dp, C.c_int, C.c_int, _dp, _dp, _dp, C.c_double, C.c_int,
                                   _i32, _dp]
        L.or_mscg_prec.restype = C.c_int
        _lib = L
    return _lib


def set_threads(t):
    lib().or_set_threads(int(t))


def threads():
    """The thread count the oracle's loops use (1 unless set_threads)."""
    return lib().or_get_threads()


def max_threads():
    return lib().or_get_max_threads()


def _csr(A):
    rp, ci, va = A
    return (np.ascontiguousarray(rp, np.int32), np.ascontiguousarray(ci, np.int32),
            np.ascontiguousarray(va, np.float64))


def spmv(A, x, alpha=1.0, beta=0.0, y=None):
    rp, ci, va = _csr(A)
    n = len(rp) - 1
    y = np.zeros(n) if y is None else np.ascontiguousarray(y, np.float64).copy()
    lib().or_spmv(n, rp, ci, va, alpha, np.ascontiguousarray(x, np.float64), beta, y)
    return y


def exp_tridiag(d, e, s, x):
    d = np.ascontiguousarray(d, np.float64)
    k = len(d)
    e = np.ascontiguousarray(np.append(e, 0.0), np.float64)
    y = np.zeros(k)
    rc = lib().or_exp_tridiag(k, d, e, s, np.ascontiguousarray(x, np.float64), y)
    if rc:
        raise RuntimeError("tridiagonal QL did not converge")
    return y


def tridiag_eig(d, e):
    d = np.ascontiguousarray(d, np.float64).copy()
    k = len(d)
    Z = np.zeros(k * k)
    rc = lib().or_tridiag_eig(k, d, np.ascontiguousarray(np.append(e, 0.0), np.float64), Z)
    if rc:
        raise RuntimeError("tridiagonal QL did not converge")
    return d, Z.reshape(k, k).T


def dense_expm(H, s=1.0):
    H = np.asfortranarray(H, dtype=np.float64)
    k = H.shape[0]
    E = np.zeros(k * k)
    rc = lib().or_dense_expm(k, np.ascontiguousarray(H.reshape(-1, order="F")), s, E)
    if rc:
        raise RuntimeError("Pade solve failed")
    return E.reshape(k, k, order="F")


def _krylov(fn, A, x, eps, itmax, istop, fscale, opscale, *extra):
    rp, ci, va = _csr(A)
    n = len(rp) - 1
    y = np.zeros(n)
    it = C.c_int(0)
    err = C.c_double(0.0)
    nres = C.c_int(0)
    res = np.zeros(max(itmax, 2) + 2)
    rc = fn(n, rp, ci, va, fscale, opscale, np.ascontiguousarray(x, np.float64), y, eps, itmax, istop,
            C.byref(it), C.byref(err), res, C.byref(nres), *extra)
    return dict(info=rc, y=y, iter=it.value, err=err.value, res=res[:nres.value].copy())


def _dfun(f):
    return DFUN(lambda z: float(f(z)))


def lanczos(A, x, eps, itmax, istop=1, fscale=1.0, opscale=1.0, f=None):
    """Reference-semantics Lanczos (krylov/polynomial/psfun_d_lanczos.f90);
    f(T) by EXP, or with f given by the USERF/SYM step with that f."""
    if f is None:
        return _krylov(lib().or_lanczos, A, x, eps, itmax, istop, fscale, opscale)
    return _krylov(lib().or_lanczos_f, A, x, eps, itmax, istop, fscale, opscale, _dfun(f))


def lanczos_batch(row_offsets, A, X, eps, itmax, istop=1, fscale=1.0, opscale=1.0):
    """or_lanczos on every (graph, column) instance of a block-diagonal batch
    (row_offsets, global-column CSR A, row-major X of nrows x K), the
    instances in parallel; returns (Y, iters (ngraph, K), errs (ngraph, K))."""
    rp, ci, va = _csr(A)
    off = np.ascontiguousarray(row_offsets, np.int64)
    ng = len(off) - 1
    X = np.ascontiguousarray(X, np.float64)
    K = 1 if X.ndim == 1 else X.shape[1]
    Y = np.zeros_like(X)
    iters = np.zeros(ng * K, np.int32)
    errs = np.zeros(ng * K)
    rc = lib().or_lanczos_batch(ng, off, rp, ci, va, K, fscale, opscale, X, Y, eps, itmax, istop, iters, errs)
    if rc:
        raise RuntimeError(f"or_lanczos_batch failed: {rc}")
    return Y, iters.reshape(ng, K), errs.reshape(ng, K)


def arnoldi(A, x, eps, itmax, istop=1, fscale=1.0, scaling=1.0, f=None):
    """Reference-semantics Arnoldi (krylov/polynomial/psfun_d_arnoldi.f90);
    with f given, f(H) by the USERF/SYM step (H's upper triangle)."""
    if f is None:
        return _krylov(lib().or_arnoldi, A, x, eps, itmax, istop, fscale, scaling)
    return _krylov(lib().or_arnoldi_f, A, x, eps, itmax, istop, fscale, scaling, _dfun(f))


def sai_arnoldi(A, x, eps, itmax, tau=0.0, istop=1, prec=0, in_itmax=1000, in_istop=2, fscale=1.0, f=None):
    """Shift-and-invert Arnoldi (krylov/rational/psfun_d_sai.f90), with the
    shift applied to every diagonal entry of H^{-1}; with f given, the
    projection by the USERF/SYM step."""
    rp, ci, va = _csr(A)
    n = len(rp) - 1
    y = np.zeros(n)
    it = C.c_int(0)
    err = C.c_double(0.0)
    nres = C.c_int(0)
    res = np.zeros(max(itmax, 2) + 2)
    args = (n, rp, ci, va, fscale, tau, prec, in_itmax, in_istop, np.ascontiguousarray(x, np.float64), y, eps,
            itmax, istop, C.byref(it), C.byref(err), res, C.byref(nres))
    if f is None:
        rc = lib().or_sai_arnoldi(*args)
    else:
        rc = lib().or_sai_arnoldi_f(*args, _dfun(f))
    return dict(info=rc, y=y, iter=it.value, err=err.value, res=res[:nres.value].copy())


def sym_eig(H):
    """(w, Z) of the symmetric matrix of H's upper triangle (DSYEV('V','U'))."""
    H = np.asfortranarray(H, dtype=np.float64)
    k = H.shape[0]
    w = np.zeros(k)
    Z = np.zeros(k * k)
    rc = lib().or_sym_eig(k, np.ascontiguousarray(H.reshape(-1, order="F")), k, w, Z)
    if rc:
        raise RuntimeError("symmetric eigen-decomposition did not converge")
    return w, Z.reshape(k, k, order="F")


def userf_sym_apply(H, s, f, x):
    """USERF/SYM f(H) step (serial/psfun_d_serial_mod.F90:420-485):
    y = Z diag(f(s*lambda)) Z^T x for the symmetric matrix of H's upper triangle."""
    H = np.asfortranarray(H, dtype=np.float64)
    k = H.shape[0]
    y = np.zeros(k)
    rc = lib().or_userf_sym_apply(k, np.ascontiguousarray(H.reshape(-1, order="F")), k, s, _dfun(f),
                                  np.ascontiguousarray(x, np.float64), y)
    if rc:
        raise RuntimeError("symmetric eigen-decomposition did not converge")
    return y


def dense_inverse(H):
    H = np.asfortranarray(H, dtype=np.float64)
    k = H.shape[0]
    out = np.zeros(k * k)
    rc = lib().or_dense_inverse(k, np.ascontiguousarray(H.reshape(-1, order="F")), out)
    if rc:
        raise RuntimeError("singular matrix")
    return out.reshape(k, k, order="F")


def ilu0(A):
    """ILU(0) factors of A (the BJAC preconditioner): (status, rowptr,
    colind, lu) of one sorted CSR holding L's strict lower part (unit
    diagonal) and U."""
    rp, ci, va = _csr(A)
    n = len(rp) - 1
    orp = np.zeros(n + 1, np.int32)
    oci = np.zeros(max(len(ci), 1), np.int32)
    olu = np.zeros(max(len(ci), 1))
    nz = C.c_int(0)
    rc = lib().or_ilu0(n, rp, ci, va, orp, oci, olu, C.byref(nz))
    return rc, orp, oci[:nz.value].copy(), olu[:nz.value].copy()


def ilu0_apply(A, r):
    """z = (LU)^{-1} r with LU the ILU(0) of A; raises on a zero pivot."""
    rp, ci, va = _csr(A)
    n = len(rp) - 1
    z = np.zeros(n)
    rc = lib().or_ilu0_apply(n, rp, ci, va, np.ascontiguousarray(r, np.float64), z)
    if rc:
        raise RuntimeError("ILU(0): zero pivot")
    return z


def cg_shifted(A, b, eta, zeta, x0=None, eps=1e-8, itmax=1000, istop=2, prec=0, want_cond=False):
    rp, ci, va = _csr(A)
    n = len(rp) - 1
    x = np.zeros(n) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
    it = C.c_int(0)
    err = C.c_double(0.0)
    cond = C.c_double(0.0)
    rc = lib().or_cg_shifted(n, rp, ci, va, prec, np.ascontiguousarray(b, np.float64), eta, zeta, x, eps,
                             itmax, istop, C.byref(it), C.byref(err), C.byref(cond) if want_cond else None)
    return dict(info=rc, x=x, iter=it.value, err=err.value, cond=cond.value if want_cond else None)


def mscg(A, b, shifts, eps=1e-8, itmax=1000):
    rp, ci, va = _csr(A)
    n = len(rp) - 1
    shifts = np.asarray(shifts, dtype=np.complex128)
    ns = len(shifts)
    sg = np.ascontiguousarray(np.stack([shifts.real, shifts.imag], 1).reshape(-1))
    X = np.zeros(2 * ns * n)
    iters = np.zeros(ns, dtype=np.int32)
    resid = np.zeros(ns)
    lib().or_mscg(n, rp, ci, va, ns, sg, np.ascontiguousarray(b, np.float64), X, eps, itmax, iters, resid)
    Xc = X.reshape(ns, n, 2)
    return dict(X=Xc[..., 0] + 1j * Xc[..., 1], iters=iters, resid=resid)


def pade_expm(H, t, degree):
    """exp(t H) by the degree-p diagonal Pade approximant with scaling and
    squaring (EXPOKIT DGPADM's method)."""
    H = np.asarray(H, np.float64)
    k = H.shape[0]
    E = np.zeros(k * k)
    rc = lib().or_pade_expm(int(degree), k, np.ascontiguousarray(H.T).reshape(-1), float(t), E)
    return rc, E.reshape(k, k).T.copy()


def set_pade_degree(p):
    """The GENPADE degree the restatements' dense exp steps use (6: exact)."""
    lib().or_set_pade_degree(int(p))


def pcocg(A, b, sigma, x0=None, prec=0, eps=1e-8, itmax=1000):
    """Preconditioned COCG on (A + sigma I) x = b (prec 0 none, 1 Jacobi of
    the shifted matrix, 2 ILU(0) of A)."""
    rp, ci, va = _csr(A)
    n = len(rp) - 1
    x = np.zeros(2 * n) if x0 is None else np.ascontiguousarray(
        np.stack([np.real(x0), np.imag(x0)], 1).reshape(-1), np.float64)
    it = C.c_int(0)
    rs = C.c_double(0.0)
    rc = lib().or_pcocg(n, rp, ci, va, prec, float(np.real(sigma)), float(np.imag(sigma)),
                        np.ascontiguousarray(b, np.float64), x, eps, itmax, C.byref(it), C.byref(rs))
    xc = x.reshape(n, 2)
    return dict(info=rc, x=xc[:, 0] + 1j * xc[:, 1], iter=it.value, resid=rs.value)


def mscg_prec(A, b, shifts, prec, eps=1e-8, itmax=1000):
    """or_mscg followed by the preconditioned per-shift COCG fallback."""
    rp, ci, va = _csr(A)
    n = len(rp) - 1
    shifts = np.asarray(shifts, dtype=np.complex128)
    ns = len(shifts)
    sg = np.ascontiguousarray(np.stack([shifts.real, shifts.imag], 1).reshape(-1))
    X = np.zeros(2 * ns * n)
    iters = np.zeros(ns, dtype=np.int32)
    resid = np.zeros(ns)
    rc = lib().or_mscg_prec(n, rp, ci, va, prec, ns, sg, np.ascontiguousarray(b, np.float64), X, eps, itmax,
                            iters, resid)
    Xc = X.reshape(ns, n, 2)
    return dict(info=rc, X=Xc[..., 0] + 1j * Xc[..., 1], iters=iters, resid=resid)


def ellipkkp(L):
    K = np.zeros(2)
    lib().or_ellipkkp(L, K)
    return K


def ellipj(u, L):
    out = np.zeros(6)
    u = complex(u)
    lib().or_ellipj(u.real, u.imag, L, out)
    return complex(out[0], out[1]), complex(out[2], out[3]), complex(out[4], out[5])


def horner(coeffs, x):
    c = np.ascontiguousarray(coeffs, np.float64)
    return lib().or_horner(len(c), c, x)


def hhtmethod1(zfun, N, rparams):
    def cb(re, im, ore, oim):
        v = complex(zfun(complex(re, im)))
        ore[0] = v.real
        oim[0] = v.imag
    cfun = ZFUN(cb)
    xi = np.zeros(2 * N)
    c = np.zeros(2 * N)
    eta = C.c_double(0)
    sign = C.c_double(0)
    lib().or_hhtmethod1(cfun, N, rparams[0], rparams[1], xi, c, C.byref(eta), C.byref(sign))
    return xi[0::2] + 1j * xi[1::2], c[0::2] + 1j * c[1::2], eta.value, sign.value


def talbot(N, t=1.0):
    s = np.zeros(2 * N)
    w = np.zeros(2 * N)
    lib().or_talbot(N, t, s, w)
    return s[0::2] + 1j * s[1::2], w[0::2] + 1j * w[1::2]


def quad_apply(A, shifts, weights, gamma, form, x, eps=1e-8, itmax=2000, prec=None):
    """prec None: the multi-shift solve alone; 0, 1, 2: with the
    preconditioned per-shift fallback (or_quad_apply_prec)."""
    rp, ci, va = _csr(A)
    n = len(rp) - 1
    shifts = np.asarray(shifts, np.complex128)
    weights = np.asarray(weights, np.complex128)
    ns = len(shifts)
    sg = np.ascontiguousarray(np.stack([shifts.real, shifts.imag], 1).reshape(-1))
    ww = np.ascontiguousarray(np.stack([weights.real, weights.imag], 1).reshape(-1))
    y = np.zeros(n)
    iters = np.zeros(ns, dtype=np.int32)
    resid = np.zeros(ns)
    rc = lib().or_quad_apply_prec(n, rp, ci, va, -1 if prec is None else int(prec), ns, sg, ww, gamma, form,
                                  np.ascontiguousarray(x, np.float64), y, eps, itmax, iters, resid)
    return dict(info=rc, y=y, iters=iters, resid=resid)


def cram_apply(H, t, y):
    """EXPOKIT DSCHBV (serial/exponential/dschbv.f): exp(t H) y for a
    symmetric H by the partial fraction form of CRAM(14,14)."""
    H = np.ascontiguousarray(np.asarray(H, np.float64).T)  # column-major H
    k = H.shape[0]
    out = np.array(y, np.float64, copy=True)
    rc = lib().or_cram_apply(k, H, k, float(t), out)
    if rc:
        raise RuntimeError(f"or_cram_apply failed: {rc}")
    return out


class exp_cram:
    """Within the block, the EXP restatements (Lanczos, Arnoldi, the batch)
    compute f(T) as the reference's CHBSYM does (DSCHBV's CRAM(14,14)),
    error floor included, instead of the exact eigen-decomposition."""

    def __enter__(self):
        lib().or_set_exp_cram(1)
        return self

    def __exit__(self, *exc):
        lib().or_set_exp_cram(0)
        return False

