"""GPU parity of the device Arnoldi f(A)x (krylov/polynomial/psfun_d_arnoldi.f90)
and of the dense f(H) variants (serial/psfun_d_serial_mod.F90:172-523)
against the CPU oracle, the DST closed form, scipy.linalg.expm on small
matrices and the reference example matrix.

Bars: y within relative 2-norm 1e-10 of the oracle (polynomial path);
iteration counts within 1; dense f(H) within 1e-12 of scipy's expm."""
import numpy as np
import pytest
import scipy.linalg as sla
import torch

from conftest import load_golden
from oracle import oracle as O
from psfun_amd import matgen, psb
from psfun_amd.arnoldi import arnoldi_apply
from psfun_amd.krylov import psfun_d_krylov
from psfun_amd.serial import EXP_VARIANTS, psfun_d_serial

pytestmark = pytest.mark.gpu


def _vec(a):
    return psb.psb_d_vect_type(torch.as_tensor(np.ascontiguousarray(a, np.float64), device="cuda:0"))


def _convdiff(m, beta=1.0):
    """Nonsymmetric 2-D convection-diffusion operator, first-order upwind in x:
    (4 + beta) on the diagonal, -(1 + beta) west, -1 east/north/south."""
    rows, cols, vals = [], [], []
    for j in range(m):
        for i in range(m):
            r = i + m * j
            for di, dj, v in ((0, 0, 4.0 + beta), (-1, 0, -1.0 - beta), (1, 0, -1.0), (0, -1, -1.0), (0, 1, -1.0)):
                ii, jj = i + di, j + dj
                if 0 <= ii < m and 0 <= jj < m:
                    rows.append(r)
                    cols.append(ii + m * jj)
                    vals.append(v)
    return matgen._coo_to_csr(m * m, np.array(rows), np.array(cols), np.array(vals))


def _arnoldi_gpu(A, x, eps, itmax, istop=1, variant="CHBHES", scaling=1.0, fmt="AUTO"):
    n = len(x)
    M = psb.psb_dspmat_type.from_csr(A)
    M.set_format(fmt)
    fun = psfun_d_serial()
    fun.set("VARIANT", variant)
    fun.set("SCALING", scaling)
    y = _vec(np.zeros(n))
    info, it, err, res = arnoldi_apply(fun, M, psb.psb_desc_type(n), y, _vec(x), eps, itmax=itmax, istop=istop)
    return info, it, err, res, y.get_vect()


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("fmt", ["CSR", "SELL", "HDIA"])
def test_arnoldi_laplacian_parity_and_closed_form(gpu_ctx, fmt):
    dims = (30, 30)
    A = matgen.laplacian(dims, scale=-1.0)
    x = matgen.uniform_pm1(31, 900)
    info, it, err, res, y = _arnoldi_gpu(A, x, 1e-10, 60, fmt=fmt)
    assert info == 0
    ref = O.arnoldi(A, x, 1e-10, 60, 1)
    assert _rel(y, ref["y"]) <= 1e-10
    assert abs(it - ref["iter"]) <= 1
    assert _rel(y, matgen.dst_expm_laplacian(dims, x, 1.0)) <= 1e-9
    assert len(res) == it + 1 and abs(res[0] - np.linalg.norm(x)) <= 1e-12 * np.linalg.norm(x)


@pytest.mark.parametrize("fmt", ["AUTO", "HDIA"])
def test_arnoldi_nonsymmetric_with_scaling(gpu_ctx, fmt):
    """SpMV multiplied by fun%scaling (arnoldi:202): exp(-0.2 A) x for a
    nonsymmetric A, against the oracle and dense scipy expm (also on the
    diagonal slices of the 5-point convection-diffusion stencil)."""
    A = _convdiff(20)
    x = matgen.normal(32, 400)
    info, it, err, res, y = _arnoldi_gpu(A, x, 1e-11, 80, scaling=-0.2, variant="GENPADE", fmt=fmt)
    assert info == 0
    ref = O.arnoldi(A, x, 1e-11, 80, 1, scaling=-0.2)
    assert _rel(y, ref["y"]) <= 1e-10
    assert abs(it - ref["iter"]) <= 1
    Ad = matgen.to_scipy(A).toarray()
    assert _rel(y, sla.expm(-0.2 * Ad) @ x) <= 1e-9


def test_arnoldi_variants_agree(gpu_ctx):
    """Every EXP variant reads the Arnoldi Hessenberg H the same way except
    CHBSYM (symmetric from the upper triangle, as DSCHBV); on a symmetric A,
    whose H is tridiagonal up to rounding, all agree."""
    A = matgen.laplacian((16, 16), scale=-1.0)
    x = matgen.uniform_pm1(33, 256)
    ys = {}
    for v in EXP_VARIANTS:
        info, it, err, res, y = _arnoldi_gpu(A, x, 1e-10, 50, variant=v)
        assert info == 0
        ys[v] = y
    ref = O.arnoldi(A, x, 1e-10, 50, 1)["y"]
    for v, y in ys.items():
        assert _rel(y, ref) <= 1e-10, v


def test_arnoldi_istop2_and_zero_vector(gpu_ctx):
    A = _convdiff(12)
    x = matgen.uniform_pm1(34, 144)
    info, it, err, res, y = _arnoldi_gpu(A, x, 1e-10, 9, istop=2, scaling=-0.1)
    ref = O.arnoldi(A, x, 1e-10, 9, 2, scaling=-0.1)
    assert info == 0 and it == 8 == ref["iter"] and err == 0.0
    assert _rel(y, ref["y"]) <= 1e-12
    info, it, err, res, y = _arnoldi_gpu(A, np.zeros(144), 1e-10, 9)
    assert info == 0 and it == 0 and np.all(y == 0.0)


def test_arnoldi_reference_example(gpu_ctx):
    """examples/polynomial/arnoldi.inp: laplacian.mtx, x = 1, EXP/GENPADE,
    scaling 1, eps 1e-6, itmax 100, istop 1."""
    g = load_golden("laplacian50.npz")
    A = (g["rowptr"] - 1, g["colind"] - 1, g["val"])
    info, it, err, res, y = _arnoldi_gpu(A, np.ones(50), 1e-6, 100, variant="GENPADE")
    ref = O.arnoldi(A, np.ones(50), 1e-6, 100, 1)
    assert info == 0 and it == ref["iter"]
    assert _rel(y, ref["y"]) <= 1e-10
    assert _rel(y, g["y_exp"]) <= 1e-6


def test_krylov_default_kname_is_arnoldi(gpu_ctx):
    A = matgen.laplacian((10, 10), scale=-1.0)
    x = matgen.uniform_pm1(35, 100)
    M = psb.psb_dspmat_type.from_csr(A)
    fun = psfun_d_serial()
    fun.set("VARIANT", "CHBHES")
    km = psfun_d_krylov()                       # kname defaults to 'ARNOLDI'
    y = _vec(np.zeros(100))
    info, it, err, res = km.apply(fun, M, psb.psb_desc_type(100), y, _vec(x), 1e-10, itmax=40)
    assert info == 0
    assert _rel(y.get_vect(), O.arnoldi(A, x, 1e-10, 40, 1)["y"]) <= 1e-10
    km.set("KNAME", "NOPE")
    assert km.apply(fun, M, psb.psb_desc_type(100), y, _vec(x), 1e-10)[0] == psb.psb_err_invalid_args_combination_


@pytest.mark.parametrize("k", [3, 12, 30])
def test_serial_dense_variants_against_scipy(gpu_ctx, k):
    g = load_golden("expm_small.npz")
    H, E = g[f"hess_{k}_H"], g[f"hess_{k}_E"]
    x = matgen.normal(40 + k, k)
    for v in ("GENPADE", "CHBGEN", "CHBHES", "TAYLOR", "SASQ"):
        fun = psfun_d_serial()
        fun.set("VARIANT", v)
        out = np.zeros(k)
        assert fun.apply(H, out, x) == 0
        assert _rel(out, E @ x) <= 1e-12, v


def test_serial_variant_matrix_models(gpu_ctx):
    """How each variant reads a full nonsymmetric H, with scaling: CHBHES its
    Hessenberg part (DNCHBV), CHBSYM the symmetric matrix of its upper
    triangle (DSCHBV, ZSYSV 'U'), the rest the whole matrix."""
    rng = np.random.default_rng(7)
    k = 20
    H = rng.normal(size=(k, k)) * 0.4
    x = rng.normal(size=k)
    s = -0.7
    up = np.triu(H)
    expect = {"GENPADE": H, "CHBGEN": H, "TAYLOR": H, "SASQ": H, "CHBHES": np.triu(H, -1),
              "CHBSYM": up + np.triu(up, 1).T}
    for v, Hm in expect.items():
        fun = psfun_d_serial()
        fun.set("VARIANT", v)
        fun.set("SCALING", s)
        out = np.zeros(k)
        assert fun.apply(H, out, x) == 0
        assert _rel(out, sla.expm(s * Hm) @ x) <= 1e-12, v


def test_arnoldi_long_run_block_estimate(gpu_ctx):
    """exp(-30 L) on 16^3 needs ~90 iterations, so the estimates past order
    64 take the block f(H) routine instead of the wavefront one.  (At eps
    1e-12 the estimate compares a component near 1e-14 of ||yk|| with eps,
    which is the rounding floor of either side's f(H); 1e-10 keeps the
    iteration counts comparable.)"""
    dims = (16, 16, 16)
    A = matgen.laplacian(dims, scale=1.0)
    x = matgen.uniform_pm1(3, 4096)
    info, it, err, res, y = _arnoldi_gpu(A, x, 1e-10, 150, scaling=-30.0, variant="GENPADE")
    assert info == 0 and it > 66
    ref = O.arnoldi(A, x, 1e-10, 150, 1, scaling=-30.0)
    assert abs(it - ref["iter"]) <= 1
    assert _rel(y, ref["y"]) <= 1e-10
    assert _rel(y, matgen.dst_expm_laplacian(dims, x, 30.0)) <= 1e-9


@pytest.mark.parametrize("k", [64, 65, 512])
def test_serial_dense_largest_orders(gpu_ctx, k):
    """The dense f(H) step at the wavefront/block boundary and at the
    largest order the C-ABI takes (512), against scipy's expm."""
    rng = np.random.default_rng(600 + k)
    H = rng.normal(size=(k, k)) * (2.0 / np.sqrt(k))
    x = rng.normal(size=k)
    for v in ("GENPADE", "CHBHES"):
        fun = psfun_d_serial()
        fun.set("VARIANT", v)
        fun.set("SCALING", -0.5)
        out = np.zeros(k)
        assert fun.apply(H, out, x) == 0
        Hm = H if v == "GENPADE" else np.triu(H, -1)
        assert _rel(out, sla.expm(-0.5 * Hm) @ x) <= 1e-12, v


def test_arnoldi_eigenvector_start_stops_at_first_estimate(gpu_ctx):
    """x an eigenvector: h(2,1) is at rounding level, so the first estimate
    (index 2) converges: iter 1, res of length 2, y = exp(-lambda) x.  The
    SpMV step of index 3 runs beside that estimate and may advance its loop
    counter; the result keeps the reference's exit index."""
    m = 20
    A = matgen.laplacian((m, m), scale=-1.0)
    i = np.arange(1, m + 1)
    p, q = 3, 5
    x = np.outer(np.sin(np.pi * i * p / (m + 1)), np.sin(np.pi * i * q / (m + 1))).reshape(-1)
    lam = (2 - 2 * np.cos(np.pi * p / (m + 1))) + (2 - 2 * np.cos(np.pi * q / (m + 1)))
    for variant in ("GENPADE", "CHBHES"):
        info, it, err, res, y = _arnoldi_gpu(A, x, 1e-10, 30, variant=variant)
        ref = O.arnoldi(A, x, 1e-10, 30, 1)
        assert info == 0 and it == 1 == ref["iter"] and len(res) == 2
        assert _rel(y, np.exp(-lam) * x) <= 1e-12
        assert _rel(y, ref["y"]) <= 1e-12


# ---- GENPADE with fun%padedegree (DGPADM's degree, serial_mod:271) ---------
@pytest.mark.parametrize("p", [2, 3, 4, 6, 10])
@pytest.mark.parametrize("k", [5, 40, 130])
def test_serial_genpade_degree(gpu_ctx, p, k):
    """fun%set('PADE_DEGREE', p): the device evaluates DGPADM's degree-p
    approximant with its scaling and squaring (degree 6: the Taylor scheme,
    equal to it at roundoff) against the oracle's restatement; a low degree
    carries its truncation error as the reference's result does."""
    rng = np.random.default_rng(700 + k)
    H = rng.normal(size=(k, k)) * (2.0 / np.sqrt(k))
    x = rng.normal(size=k)
    fun = psfun_d_serial()
    fun.set("VARIANT", "GENPADE")
    fun.set("SCALING", -0.7)
    fun.set("PADE_DEGREE", p)
    out = np.zeros(k)
    assert fun.apply(H, out, x) == 0
    rc, E = O.pade_expm(H, -0.7, p)
    assert rc == 0 and _rel(out, E @ x) <= 1e-12
    if p == 2 and k == 40:
        assert _rel(out, sla.expm(-0.7 * H) @ x) > 1e-9     # 5.3e-9 here


def test_serial_genpade_degree_limits(gpu_ctx):
    fun = psfun_d_serial()
    fun.set("VARIANT", "GENPADE")
    out = np.zeros(3)
    for p in (1, 31):
        fun.set("PADE_DEGREE", p)
        assert fun.apply(np.eye(3), out, np.ones(3)) != 0
    fun.set("PADE_DEGREE", 3)
    assert fun.apply(np.zeros((3, 3)), out, np.ones(3)) == 0 and np.all(out == 1.0)   # exp(0) = I


@pytest.mark.parametrize("p", [2, 3])
def test_arnoldi_and_sai_genpade_degree(gpu_ctx, p):
    """Arnoldi and shift-and-invert Arnoldi with GENPADE at degree p: every
    f(H) step (the estimates and the final projection) uses the approximant,
    against the oracle run at the same degree; degree 2 leaves 1e-6 of
    truncation in exp(-0.2 A)x, as the reference would."""
    A = _convdiff(20)
    x = matgen.normal(32, 400)
    M = psb.psb_dspmat_type.from_csr(A)
    fun = psfun_d_serial()
    fun.set("VARIANT", "GENPADE")
    fun.set("SCALING", -0.2)
    fun.set("PADE_DEGREE", p)
    y = _vec(np.zeros(400))
    info, it, err, res = arnoldi_apply(fun, M, psb.psb_desc_type(400), y, _vec(x), 1e-11, itmax=80, istop=1)
    O.set_pade_degree(p)
    try:
        ref = O.arnoldi(A, x, 1e-11, 80, 1, scaling=-0.2)
    finally:
        O.set_pade_degree(6)
    assert info == 0 and abs(it - ref["iter"]) <= 1
    assert _rel(y.get_vect(), ref["y"]) <= 1e-10
    ex = sla.expm(-0.2 * matgen.to_scipy(A).toarray()) @ x
    if p == 2:
        assert _rel(y.get_vect(), ex) > 1e-7
    # SAI on the Laplacian (fscale carries the time)
    from psfun_amd.arnoldi import saiarnoldi_apply
    L = matgen.laplacian((16, 16), scale=-1.0)
    xs = matgen.normal(33, 256)
    Ms = psb.psb_dspmat_type.from_csr(L)
    fun2 = psfun_d_serial()
    fun2.set("VARIANT", "GENPADE")
    fun2.set("PADE_DEGREE", p)
    ys = _vec(np.zeros(256))
    info, it, err, res = saiarnoldi_apply(fun2, Ms, "CG", None, -1.0, psb.psb_desc_type(256), 4000, 0, 2, ys,
                                          _vec(xs), 1e-9, itmax=60, istop=1)
    O.set_pade_degree(p)
    try:
        ref = O.sai_arnoldi(L, xs, 1e-9, 60, tau=-1.0, istop=1, in_itmax=4000, in_istop=2)
    finally:
        O.set_pade_degree(6)
    assert info == 0 and abs(it - ref["iter"]) <= 1
    assert _rel(ys.get_vect(), ref["y"]) <= 1e-8
