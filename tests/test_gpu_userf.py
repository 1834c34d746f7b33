"""GPU parity of the USERF/SYM f(H) step (serial/psfun_d_serial_mod.F90:420-485:
DSYEV('V','U') + f_ptr + DGEMV) as a serial apply and inside the Lanczos,
Arnoldi and shift-and-invert Arnoldi loops, against the oracle restatement
(Householder + implicit QL, itself pinned to LAPACK eigh) and scipy's expm /
sqrtm.  The device diagonalises by block Jacobi, so the two sides agree to
rounding: bars 1e-12 for the f(H) step, 1e-10 for the polynomial Krylov
path, 1e-8 for shift-and-invert (inner CG to the outer tolerance)."""
import math

import numpy as np
import pytest
import scipy.linalg as sla
import torch

from oracle import oracle as O
from psfun_amd import matgen, psb
from psfun_amd.arnoldi import arnoldi_apply, saiarnoldi_apply
from psfun_amd.krylov import lanczos_apply
from psfun_amd.serial import psfun_d_serial

pytestmark = pytest.mark.gpu


def _vec(a):
    return psb.psb_d_vect_type(torch.as_tensor(np.ascontiguousarray(a, np.float64), device="cuda:0"))


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _userf(f, scaling=1.0):
    fun = psfun_d_serial()
    fun.set("FNAME", "USERF")
    fun.set("VARIANT", "SYM")
    fun.set("FPOINTER", f)
    fun.set("SCALING", scaling)
    return fun


@pytest.mark.parametrize("k", [1, 2, 7, 40, 131])
def test_serial_userf_against_oracle_and_scipy(gpu_ctx, k):
    rng = np.random.default_rng(500 + k)
    H = rng.normal(size=(k, k))
    S = np.triu(H) + np.triu(H, 1).T          # what DSYEV('V','U') sees
    x = rng.normal(size=k)
    out = np.zeros(k)
    assert _userf(np.exp, -0.3).apply(H, out, x) == 0
    assert _rel(out, O.userf_sym_apply(H, -0.3, np.exp, x)) <= 1e-12
    assert _rel(out, sla.expm(-0.3 * S) @ x) <= 1e-12
    B = rng.normal(size=(k, k))
    P = B @ B.T + k * np.eye(k)
    assert _userf(np.sqrt).apply(P, out, x) == 0
    assert _rel(out, np.real(sla.sqrtm(P)) @ x) <= 1e-12


def test_serial_userf_sparse_matrix_and_errors(gpu_ctx):
    """A sparse matrix is densified first (psfun_d_serial_apply_sparse);
    USERF without a function or with another variant fails like the
    reference's case default; an exception in f propagates."""
    A = matgen.laplacian((5, 5), scale=-1.0)
    x = matgen.uniform_pm1(9, 25)
    M = psb.psb_dspmat_type.from_csr(A)
    out = np.zeros(25)
    assert _userf(np.exp).apply(M, out, x) == 0
    assert _rel(out, sla.expm(matgen.to_scipy(A).toarray()) @ x) <= 1e-12
    fun = psfun_d_serial()
    fun.set("FNAME", "USERF")
    fun.set("VARIANT", "SYM")
    assert fun.apply(np.eye(3), np.zeros(3), np.ones(3)) == psb.psb_err_from_subroutine_   # no f_ptr
    fun = _userf(np.exp)
    fun.set("VARIANT", "CHBSYM")
    assert fun.apply(np.eye(3), np.zeros(3), np.ones(3)) == psb.psb_err_from_subroutine_
    fun = psfun_d_serial()
    fun.set("FNAME", "PHI")
    fun.set("VARIANT", "SYM")
    assert fun.apply(np.eye(3), np.zeros(3), np.ones(3)) == psb.psb_err_from_subroutine_

    def bad(z):
        raise ValueError("user function failed")

    with pytest.raises(ValueError):
        _userf(bad).apply(np.eye(3), np.zeros(3), np.ones(3))


@pytest.mark.parametrize("istop", [1, 2])
def test_lanczos_userf_exp_parity(gpu_ctx, istop):
    """f = exp through the user function: same iterations and result as the
    oracle's USERF Lanczos and as the device EXP path."""
    dims = (24, 24)
    A = matgen.laplacian(dims, scale=-1.0)
    n = 576
    x = matgen.uniform_pm1(41, n)
    M = psb.psb_dspmat_type.from_csr(A)
    y = _vec(np.zeros(n))
    itmax = 60 if istop == 1 else 25
    info, it, err, res = lanczos_apply(_userf(np.exp), M, psb.psb_desc_type(n), y, _vec(x), 1e-10, itmax=itmax,
                                       istop=istop)
    assert info == 0
    ref = O.lanczos(A, x, 1e-10, itmax, istop, f=np.exp)
    assert abs(it - ref["iter"]) <= 1
    assert _rel(y.get_vect(), ref["y"]) <= 1e-10
    m = min(len(res), len(ref["res"])) - 1          # estimates rs(1:itx), same on both sides
    assert np.allclose(res[:m], ref["res"][:m], rtol=1e-6, atol=1e-13 * res[0])
    fexp = psfun_d_serial()
    fexp.set("VARIANT", "CHBSYM")
    y2 = _vec(np.zeros(n))
    info, it2, _, _ = lanczos_apply(fexp, M, psb.psb_desc_type(n), y2, _vec(x), 1e-10, itmax=itmax, istop=istop)
    assert info == 0 and abs(it2 - it) <= 1
    assert _rel(y.get_vect(), y2.get_vect()) <= 1e-10
    if istop == 1:
        assert _rel(y.get_vect(), matgen.dst_expm_laplacian(dims, x, 1.0)) <= 1e-9


def test_lanczos_userf_sqrt(gpu_ctx):
    """f = sqrt of the SPD Laplacian (opscale +1): oracle parity and sqrtm.
    The user function is math.sqrt, which raises on a negative argument, and
    only ever sees eigenvalues of the reference's T_k (inside A's spectrum):
    steps after the loop has stopped never call it."""
    dims = (9, 9)
    A = matgen.laplacian(dims, scale=1.0)
    x = matgen.uniform_pm1(42, 81)
    M = psb.psb_dspmat_type.from_csr(A)
    y = _vec(np.zeros(81))
    seen = []

    def f(z):
        seen.append(z)
        return math.sqrt(z)

    info, it, err, res = lanczos_apply(_userf(f), M, psb.psb_desc_type(81), y, _vec(x), 1e-12, itmax=81)
    assert info == 0
    lam = np.linalg.eigvalsh(matgen.to_scipy(A).toarray())
    assert min(seen) >= lam[0] * (1 - 1e-9) and max(seen) <= lam[-1] * (1 + 1e-9)
    ref = O.lanczos(A, x, 1e-12, 81, 1, f=np.sqrt)
    assert abs(it - ref["iter"]) <= 1
    assert _rel(y.get_vect(), ref["y"]) <= 1e-10
    Ad = matgen.to_scipy(A).toarray()
    assert _rel(y.get_vect(), np.real(sla.sqrtm(Ad)) @ x) <= 1e-9


def test_arnoldi_userf_parity(gpu_ctx):
    """USERF/SYM reads H_k's upper triangle; with scaling in the SpMV."""
    A = matgen.laplacian((16, 16), scale=1.0)
    x = matgen.uniform_pm1(43, 256)
    M = psb.psb_dspmat_type.from_csr(A)
    y = _vec(np.zeros(256))
    fun = _userf(np.exp, scaling=-0.5)
    info, it, err, res = arnoldi_apply(fun, M, psb.psb_desc_type(256), y, _vec(x), 1e-10, itmax=60)
    assert info == 0
    ref = O.arnoldi(A, x, 1e-10, 60, 1, scaling=-0.5, f=np.exp)
    assert abs(it - ref["iter"]) <= 1
    assert _rel(y.get_vect(), ref["y"]) <= 1e-10
    assert _rel(y.get_vect(), matgen.dst_expm_laplacian((16, 16), x, 0.5)) <= 1e-9


def test_sai_userf_parity(gpu_ctx):
    dims = (12, 12)
    A = matgen.laplacian(dims, scale=-10.0)
    n = 144
    x = matgen.uniform_pm1(44, n)
    M = psb.psb_dspmat_type.from_csr(A)
    y = _vec(np.zeros(n))
    info, it, err, res = saiarnoldi_apply(_userf(np.exp), M, "CG", None, 0.0, psb.psb_desc_type(n), 2000, 0, 2, y,
                                          _vec(x), 1e-10, itmax=60, istop=1)
    assert info == 0
    ref = O.sai_arnoldi(A, x, 1e-10, 60, tau=0.0, istop=1, in_itmax=2000, in_istop=2, f=np.exp)
    assert abs(it - ref["iter"]) <= 1
    assert _rel(y.get_vect(), ref["y"]) <= 1e-8
    assert _rel(y.get_vect(), matgen.dst_expm_laplacian(dims, x, 10.0)) <= 1e-8
