"""Pins the CPU oracle (oracle/psfun_oracle.c) before it is trusted as the
checker: reference known-answer values, closed forms, scipy dense references
and brute force, each chosen so that a dropped term or wrong index fails."""
import numpy as np
import pytest
import scipy.linalg as sla

from conftest import load_golden
from oracle import oracle as O
from psfun_amd import matgen


# ---- examples/utils/utiltest.f90:63-87 control values ----------------------
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-0.5, 2.0), (3.0, -1.0), (0.0, 0.5)])
def test_spmv_against_scipy(alpha, beta):
    """or_spmv (y = alpha A x + beta y) against scipy's CSR product on a ragged
    matrix with empty rows, unsorted columns and repeated (row, col) entries,
    which a CSR product sums."""
    import scipy.sparse as sps
    rng = np.random.default_rng(11)
    n, nnz = 57, 400
    rows = np.sort(rng.integers(0, n, nnz))
    rows[rows == 7] = 8                      # row 7 stays empty
    cols = rng.integers(0, n, nnz)
    cols[:40] = cols[40:80]                  # repeat some columns ...
    rows[:40] = rows[40:80]                  # ... in the same rows
    order = np.argsort(rows, kind="stable")
    rows, cols = rows[order], cols[order]
    vals = rng.standard_normal(nnz)
    rp = np.zeros(n + 1, np.int32)
    np.cumsum(np.bincount(rows, minlength=n), out=rp[1:])
    S = sps.csr_matrix((vals, cols.astype(np.int32), rp), shape=(n, n))
    assert not S.has_canonical_format        # duplicates are present
    x = rng.standard_normal(n)
    y0 = rng.standard_normal(n)
    got = O.spmv((rp, cols, vals), x, alpha=alpha, beta=beta, y=y0)
    want = alpha * (S @ x) + beta * y0
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)
    assert got[7] == beta * y0[7]


def test_horner_kat():
    assert O.horner([132.0, 42.0, 14.0, 5.0, 2.0, 1.0, 0.0], 1.0) == 196.0
    # brute-force polynomial on another point
    c = np.array([3.0, -1.0, 0.5, 2.0])
    x = 1.7
    assert abs(O.horner(c, x) - sum(c[i] * x ** i for i in range(4))) < 1e-13


def test_ellipkkp_kat():
    K = O.ellipkkp(0.5)
    assert abs(K[0] - 1.588191701877384) < 1e-14
    assert abs(K[1] - 2.978718299395645) < 1e-14


def test_ellipkkp_matches_scipy():
    from scipy.special import ellipk, ellipkm1
    for L in (0.05, 0.3, 1.0, 3.0, 9.5):
        m = np.exp(-2 * np.pi * L)
        K = O.ellipkkp(L)
        assert abs(K[0] - ellipk(m)) < 1e-13 * ellipk(m)
        # K(1-m) for tiny m: ellipkm1 takes m itself (no cancellation in 1-m)
        assert abs(K[1] - ellipkm1(m)) < 1e-12 * ellipkm1(m)
    # L > 10 asymptotic branch
    K = O.ellipkkp(12.0)
    assert K[0] == np.pi / 2 and abs(K[1] - (np.pi * 12 + np.log(4))) < 1e-14


def test_ellipj_real_kat():
    sn, cn, dn = O.ellipj(1.0, 0.5)
    assert abs(sn - 0.838274911024583) < 1e-14
    assert abs(cn - 0.545247809300255) < 1e-14
    assert abs(dn - 0.984699634947678) < 1e-14


def test_ellipj_complex_kat():
    sn, cn, dn = O.ellipj(0.5j, 1.0)
    assert abs(sn - 0.521141424070547j) < 1e-14
    assert abs(cn - 1.127647278133671) < 1e-14
    assert abs(dn - 1.000253555731494) < 1e-14


def test_ellipj_real_matches_scipy():
    from scipy.special import ellipj
    for L in (0.1, 0.5, 2.0):
        m = np.exp(-2 * np.pi * L)
        for u in (0.1, 0.7, 1.3):
            sn, cn, dn = O.ellipj(u, L)
            s2, c2, d2, _ = ellipj(u, m)
            assert abs(sn - s2) < 1e-13 and abs(cn - c2) < 1e-13 and abs(dn - d2) < 1e-13


def test_ellipj_complex_identities():
    # sn^2 + cn^2 = 1 and dn^2 + m sn^2 = 1 hold for complex arguments too
    for L in (0.2, 0.8):
        m = np.exp(-2 * np.pi * L)
        for u in (0.3 + 0.2j, -0.9 + 0.5j, 1.1 + 0.05j):
            sn, cn, dn = O.ellipj(u, L)
            assert abs(sn * sn + cn * cn - 1) < 1e-12
            assert abs(dn * dn + m * sn * sn - 1) < 1e-12


# ---- small f(T) step -------------------------------------------------------
def test_exp_tridiag_against_scipy_fixtures():
    g = load_golden("expm_small.npz")
    for k in (1, 2, 5, 17, 40, 64, 65, 130):
        for s in (1.0, -0.5):
            d, e, y = g[f"tri_{k}_{s}_d"], g[f"tri_{k}_{s}_e"], g[f"tri_{k}_{s}_y"]
            e1 = np.zeros(k)
            e1[0] = 1
            got = O.exp_tridiag(d, e, s, e1)
            assert np.linalg.norm(got - y) <= 1e-12 * np.linalg.norm(y)
            # the trailing component drives the stopping test: the eigen-decomposition
            # gets it to an absolute (not relative) accuracy of a few ulps of ||y||
            assert abs(got[-1] - y[-1]) <= 1e-14 * np.linalg.norm(y), (k, s, got[-1], y[-1])


def test_exp_tridiag_2x2_closed_form():
    a, b, c = -1.0, 0.3, -2.0
    lam = np.linalg.eigvalsh(np.array([[a, b], [b, c]]))
    d, e = np.array([a, c]), np.array([b])
    y = O.exp_tridiag(d, e, 1.0, np.array([1.0, 0.0]))
    # exp(T) e1 via the spectral formula for 2x2 symmetric matrices
    l1, l2 = lam
    E = (np.exp(l1) * (np.array([[a, b], [b, c]]) - l2 * np.eye(2)) -
         np.exp(l2) * (np.array([[a, b], [b, c]]) - l1 * np.eye(2))) / (l1 - l2)
    assert np.allclose(y, E[:, 0], rtol=1e-14, atol=1e-15)


def test_cram_scalar_error_curve():
    """DSCHBV's CRAM(14,14) on 1x1 matrices: r(x) - exp(-x) for x >= 0 is the
    approximation's own error.  A wrong digit or sign in any pole or residue
    moves it far above the table's level (max |error| 3.13e-12, at x = 0, the
    value EXPOKIT's coefficients give in exact arithmetic), and at infinity
    the error is r(inf) = alpha0 = 1.83216998528140087e-12 (dschbv.f:56)."""
    xs = np.concatenate([np.linspace(0.0, 2.0, 401), np.linspace(2.0, 60.0, 581), [1e3, 1e6]])
    err = np.array([O.cram_apply(np.array([[-x]]), 1.0, np.array([1.0]))[0] - np.exp(-x) for x in xs])
    assert np.abs(err).max() <= 3.2e-12
    assert abs(err[0] + 3.13e-12) <= 0.01e-12          # the largest deviation, at x = 0
    assert abs(err[-1] - 1.83216998528140087e-12) <= 1e-15
    assert np.any(err > 5e-13) and np.any(err < -5e-13)  # an error curve that changes sign, not a bias
    # t scales H (exp(t H)): the same value for H = -x/2 at t = 2
    assert O.cram_apply(np.array([[-0.25]]), 2.0, np.array([1.0]))[0] == pytest.approx(np.exp(-0.5), abs=3.2e-12)


def test_cram_dense_against_scipy():
    """DSCHBV on dense symmetric negative definite matrices: within the
    approximation's error of scipy's expm, and the exact eigen-decomposition
    path of the same oracle is 1e-3 of that away."""
    rng = np.random.default_rng(4)
    for k in (1, 5, 30):
        B = rng.standard_normal((k, k))
        H = -(B @ B.T) / k - 0.1 * np.eye(k)
        y = rng.standard_normal(k)
        ref = sla.expm(0.7 * H) @ y
        got = O.cram_apply(H, 0.7, y)
        assert np.linalg.norm(got - ref) <= 4e-12 * np.linalg.norm(y)


@pytest.mark.parametrize("dims,itmax,it_expected", [((100, 100), 30, 20), ((128, 128, 128), 50, 25)])
def test_lanczos_cram_trajectory_matches_exact(dims, itmax, it_expected):
    """The reference's CHBSYM computes f(T) by CRAM(14,14) (about 3e-12
    absolute error); this oracle's default and the device compute it to
    rounding.  On the workloads of the CPU-runnable config (100^2, t = 1,
    itmax 30) and the plain GPU config (128^3, t = 1, itmax 50) at eps 1e-10
    both stop at the same index, their error estimates agree to 1e-6 while
    above 1e-2 and to 1e-2 down to the last two (CRAM's error reaches the
    small trailing components), and the reference's y carries CRAM's own
    error (a few 1e-12 of the closed form) where the exact f(T) gives ~1e-14
    at 128^3."""
    A = matgen.laplacian(dims)
    n = len(A[0]) - 1
    x = matgen.uniform_pm1(1, n)
    r0 = O.lanczos(A, x, 1e-10, itmax, 1, opscale=-1.0)
    with O.exp_cram():
        r1 = O.lanczos(A, x, 1e-10, itmax, 1, opscale=-1.0)
    ex = matgen.dst_expm_laplacian(dims, x, 1.0)
    assert r0["iter"] == r1["iter"] == it_expected
    assert r1["err"] < 1e-10
    k = r0["iter"]
    a, b = r0["res"][1:k - 2], r1["res"][1:k - 2]  # all but the last two estimates
    assert np.allclose(b, a, rtol=1e-2, atol=0)
    assert np.allclose(b[a > 1e-2], a[a > 1e-2], rtol=1e-6, atol=0)
    e0 = np.linalg.norm(r0["y"] - ex) / np.linalg.norm(ex)
    e1 = np.linalg.norm(r1["y"] - ex) / np.linalg.norm(ex)
    assert e0 <= 1e-12 and 5e-13 <= e1 <= 1e-11


def test_dense_expm_against_scipy():
    g = load_golden("expm_small.npz")
    for k in (3, 12, 30):
        E = O.dense_expm(g[f"hess_{k}_H"])
        assert np.linalg.norm(E - g[f"hess_{k}_E"]) <= 1e-12 * np.linalg.norm(g[f"hess_{k}_E"])
    # nilpotent: exp(N) = I + N + N^2/2
    N = np.diag([1.0, 2.0], 1)
    assert np.allclose(O.dense_expm(N), np.eye(3) + N + N @ N / 2, rtol=0, atol=1e-15)


# ---- Lanczos / Arnoldi ------------------------------------------------------
def _dirichlet_case(dims, seed, t=1.0):
    A = matgen.laplacian(dims, scale=-t)   # the reference ignores fun%scaling: pass -tA
    x = matgen.uniform_pm1(seed, int(np.prod(dims)))
    return A, x, matgen.dst_expm_laplacian(dims, x, t)


@pytest.mark.parametrize("dims", [(100, 100), (12, 10, 9)])
def test_lanczos_dst_closed_form(dims):
    A, x, yex = _dirichlet_case(dims, 1)
    r = O.lanczos(A, x, 1e-10, 30 if len(dims) == 2 else 50, 1)
    assert r["info"] == 0
    rel = np.linalg.norm(r["y"] - yex) / np.linalg.norm(yex)
    assert rel < 1e-10
    # res(1) = ||x||, res(itx+1) = err < eps, err below eps on convergence
    assert abs(r["res"][0] - np.linalg.norm(x)) < 1e-12 * np.linalg.norm(x)
    assert len(r["res"]) == r["iter"] + 1 and r["res"][-1] == r["err"] < 1e-10


def test_lanczos_reference_example_laplacian50():
    g = load_golden("laplacian50.npz")
    A = (g["rowptr"] - 1, g["colind"] - 1, g["val"])
    r = O.lanczos(A, np.ones(50), 1e-6, 100, 1)   # lanczostest.F90:213-216 overrides
    rel = np.linalg.norm(r["y"] - g["y_exp"]) / np.linalg.norm(g["y_exp"])
    assert rel < 1e-6 and r["err"] < 1e-6


def test_lanczos_matches_dense_expm_with_fscale():
    A = matgen.laplacian((7, 6))
    Ad = matgen.to_scipy(A).toarray()
    x = matgen.uniform_pm1(3, 42)
    r = O.lanczos(A, x, 1e-12, 40, 1, fscale=-0.3)
    y = sla.expm(-0.3 * Ad) @ x
    assert np.linalg.norm(r["y"] - y) < 1e-11 * np.linalg.norm(y)


def test_lanczos_maxit_drops_to_nl_minus_2():
    # istop = 2: no estimate, the result uses T(1:nl-2) and nl-2 basis vectors
    A, x, _ = _dirichlet_case((20, 20), 5)
    nl = 8
    r = O.lanczos(A, x, 1e-10, nl, 2)
    assert r["iter"] == nl - 1 and r["err"] == 0.0
    # brute force: build the basis by dense Lanczos and assemble with k = nl-2
    Ad = matgen.to_scipy(A).toarray()
    V = np.zeros((400, nl))
    V[:, 0] = x / np.linalg.norm(x)
    al, be = [], []
    w = Ad @ V[:, 0]
    al.append(w @ V[:, 0])
    w = w - al[0] * V[:, 0]
    for i in range(1, nl - 1):
        be.append(np.linalg.norm(w))
        V[:, i] = w / be[-1]
        w = Ad @ V[:, i]
        al.append(w @ V[:, i])
        w = w - al[-1] * V[:, i] - be[-1] * V[:, i - 1]
    k = nl - 2
    T = np.diag(al[:k]) + np.diag(be[:k - 1], 1) + np.diag(be[:k - 1], -1)
    yk = sla.expm(T)[:, 0]
    y = np.linalg.norm(x) * V[:, :k] @ yk
    assert np.linalg.norm(r["y"] - y) < 1e-12 * np.linalg.norm(y)


def test_lanczos_breakdown_on_eigenvector():
    # x = eigenvector of the 1-D Laplacian -> beta_1 = 0 at the first check
    m = 30
    A = matgen.laplacian((m,), scale=-1.0)
    k = 3
    x = np.sin(np.arange(1, m + 1) * k * np.pi / (m + 1))
    lam = -(2 - 2 * np.cos(k * np.pi / (m + 1)))
    r = O.lanczos(A, x, 1e-10, 20, 1)
    assert r["iter"] == 1
    assert np.linalg.norm(r["y"] - np.exp(lam) * x) < 1e-12 * np.linalg.norm(x)


def test_lanczos_invalid_istop():
    A = matgen.laplacian((5,))
    assert O.lanczos(A, np.ones(5), 1e-8, 10, 3)["info"] == 2


def test_lanczos_batch_is_or_lanczos_per_instance():
    """The batched restatement runs or_lanczos on every (graph, column)
    instance: bitwise the single-threaded per-instance result, also for a
    1-node graph, an eigenvector column (breakdown) and a zero column (NaN,
    as the reference's 1/||x|| gives)."""
    from psfun_amd.batched import concat_block_diagonal
    sizes = [1000, 1, 50, 30, 2000]
    mats = [matgen.rgg_laplacian(n, 300 + i) for i, n in enumerate(sizes)]
    mats[3] = matgen.laplacian((30,))
    off, rp, ci, va = concat_block_diagonal(mats)
    K = 3
    X = matgen.normal(8, sum(sizes) * K).reshape(sum(sizes), K)
    X[off[3]:off[4], 1] = np.sin(np.arange(1, 31) * 2 * np.pi / 31)
    X[off[2]:off[3], 2] = 0.0
    saved = O.threads()
    try:
        O.set_threads(4)
        Y, it, er = O.lanczos_batch(off, (rp, ci, va), X, 1e-10, 40, 1, opscale=-0.5)
        O.set_threads(1)
        for g, A in enumerate(mats):
            for k in range(K):
                r = O.lanczos(A, X[off[g]:off[g + 1], k], 1e-10, 40, 1, opscale=-0.5)
                assert np.array_equal(r["y"], Y[off[g]:off[g + 1], k], equal_nan=True), (g, k)
                assert r["iter"] == it[g, k]
                assert r["err"] == er[g, k] or (np.isnan(r["err"]) and np.isnan(er[g, k]))
    finally:
        O.set_threads(saved)
    assert it[3, 1] == 1


def test_arnoldi_dst_closed_form():
    A, x, yex = _dirichlet_case((30, 30), 2)
    r = O.arnoldi(A, x, 1e-10, 40, 1)
    assert np.linalg.norm(r["y"] - yex) / np.linalg.norm(yex) < 1e-10


def test_arnoldi_uses_scaling_in_spmv():
    A = matgen.laplacian((30, 30))
    x = matgen.uniform_pm1(2, 900)
    yex = matgen.dst_expm_laplacian((30, 30), x, 0.5)
    r = O.arnoldi(A, x, 1e-10, 40, 1, scaling=-0.5)
    assert np.linalg.norm(r["y"] - yex) / np.linalg.norm(yex) < 1e-10


# ---- shifted CG -----------------------------------------------------------
def test_cg_shifted_reference_example():
    g = load_golden("laplacian50.npz")
    A = (g["rowptr"] - 1, g["colind"] - 1, g["val"])
    # shifted.inp: eta=-1, zeta=1, DIAG, istop 2, itmax 500, eps 1e-7
    r = O.cg_shifted(A, np.ones(50), -1.0, 1.0, eps=1e-7, itmax=500, istop=2, prec=1)
    assert np.linalg.norm(r["x"] - g["x_shift"]) < 1e-6 * np.linalg.norm(g["x_shift"])
    assert r["err"] <= 1e-7


@pytest.mark.parametrize("istop", [1, 2])
def test_cg_shifted_dense_solve(istop):
    A = matgen.laplacian((9, 8))
    Ad = matgen.to_scipy(A).toarray()
    b = matgen.uniform_pm1(4, 72)
    eta, zeta = 0.7, 0.3
    r = O.cg_shifted(A, b, eta, zeta, eps=1e-12, itmax=500, istop=istop, want_cond=True)
    x = np.linalg.solve(eta * Ad + zeta * np.eye(72), b)
    assert np.linalg.norm(r["x"] - x) < 1e-9 * np.linalg.norm(x)
    ev = np.linalg.eigvalsh(eta * Ad + zeta * np.eye(72))
    assert abs(r["cond"] - ev[-1] / ev[0]) < 1e-6 * ev[-1] / ev[0]


def test_cg_itmax_stops():
    A = matgen.laplacian((20, 20))
    r = O.cg_shifted(A, np.ones(400), 1.0, 0.0, eps=1e-30, itmax=7, istop=2)
    assert r["iter"] == 7


# Breakdowns with an indefinite DIAG preconditioner (psfun_dcg.F90:280-310),
# traced by hand.  b = (1, 1), dinv = (1, -1) gives rho = r.z = 0 at it = 1.
#  * A = [[1, .5], [.5, -1]]: sigma = p.Ap = -1, alpha = 0 (x, r unchanged);
#    it = 2 has rho_old = 0: breakdown, counted (itx = itx + 1 comes first);
#    so every restart costs 2 iterations, and a check at itx = itmax ends it.
#  * A = diag(1, -1): p = z = (1, -1), sigma = p.Ap = 0 at it = 1: each
#    restart costs one iteration until itx = itmax.
CG_RHO_BRK = (np.array([0, 2, 4], np.int32), np.array([0, 1, 0, 1], np.int32), np.array([1.0, 0.5, 0.5, -1.0]))
CG_SIG_BRK = (np.array([0, 1, 2], np.int32), np.array([0, 1], np.int32), np.array([1.0, -1.0]))


@pytest.mark.parametrize("itmax", [1, 2, 7, 8])
def test_cg_breakdown_restarts_count_iterations(itmax):
    r = O.cg_shifted(CG_RHO_BRK, np.ones(2), 1.0, 0.0, eps=1e-8, itmax=itmax, istop=2, prec=1)
    assert r["iter"] == itmax and np.all(r["x"] == 0.0)
    assert r["err"] == 1.0       # ||r|| / ||b|| of the last check (x = 0)
    r = O.cg_shifted(CG_SIG_BRK, np.ones(2), 1.0, 0.0, eps=1e-8, itmax=itmax, istop=2, prec=1)
    assert r["iter"] == itmax and np.all(r["x"] == 0.0)
    assert r["err"] == 0.0       # no convergence check ran (psb_init_conv's errnum = 0)


# ---- ILU(0), the BJAC preconditioner with one process ----------------------
def _from_dense(M):
    import scipy.sparse as sps
    S = sps.csr_matrix(M)
    S.sort_indices()
    return S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data.astype(np.float64)


def _random_sparse(n, density, seed, diag_shift):
    import scipy.sparse as sps
    S = sps.random(n, n, density=density, random_state=seed, format="csr") + diag_shift * sps.eye(n, format="csr")
    return _from_dense(S.toarray())


def _ilu_factors(rp, ci, lu):
    n = len(rp) - 1
    L, U = np.eye(n), np.zeros((n, n))
    for i in range(n):
        for q in range(rp[i], rp[i + 1]):
            (L if ci[q] < i else U)[i, ci[q]] = lu[q]
    return L, U


@pytest.mark.parametrize("make", [lambda: matgen.anisotropic_2d(7, seed=5), lambda: matgen.laplacian((5, 4, 3)),
                                  lambda: _random_sparse(40, 0.12, 2, 6.0)])
def test_ilu0_product_equals_a_on_pattern(make):
    """ILU(0) is fixed by two conditions: L and U live on A's pattern and
    (L U)_ij = a_ij on that pattern; a wrong update order or a missed term
    breaks the second."""
    A = make()
    S = matgen.to_scipy(A).toarray()
    rc, rp, ci, lu = O.ilu0(A)
    assert rc == 0
    L, U = _ilu_factors(rp, ci, lu)
    pat = S != 0
    assert np.array_equal((L != 0) | (U != 0) | np.eye(len(S), dtype=bool), pat | np.eye(len(S), dtype=bool))
    P = L @ U
    assert np.max(np.abs(P[pat] - S[pat])) <= 1e-13 * np.max(np.abs(S))
    # fill outside the pattern is dropped, so M differs from A there (for a 2-D stencil)
    if not np.allclose(P, S):
        assert np.max(np.abs(P[~pat])) > 0


@pytest.mark.parametrize("n", [1, 2, 17])
def test_ilu0_tridiagonal_and_dense_are_exact(n):
    """No fill for a tridiagonal or a dense pattern: M = A, so M^{-1} A x = x."""
    rng = np.random.default_rng(n)
    T = np.diag(4.0 + rng.random(n)) + np.diag(rng.random(n - 1) - 0.5, 1) + np.diag(rng.random(n - 1) - 0.5, -1)
    D = rng.random((n, n)) + n * np.eye(n)
    for M in (T, D):
        sp = _from_dense(M)
        x = rng.standard_normal(n)
        z = O.ilu0_apply(sp, M @ x)
        assert np.linalg.norm(z - x) <= 1e-13 * np.linalg.norm(x) * n
    # the dense pattern's factors are the unpivoted LU: L U = D everywhere
    rc, rp, ci, lu = O.ilu0(_from_dense(D))
    L, U = _ilu_factors(rp, ci, lu)
    assert rc == 0 and np.max(np.abs(L @ U - D)) <= 1e-13 * n * np.max(np.abs(D))


def test_ilu0_zero_pivot_breakdown():
    """A zero leading entry (or a row without a diagonal) stops the build."""
    Z = _from_dense(np.array([[0.0, 1.0], [1.0, 2.0]]))
    assert O.ilu0(Z)[0] == 10
    with pytest.raises(RuntimeError):
        O.ilu0_apply(Z, np.ones(2))
    assert O.cg_shifted(Z, np.ones(2), 1.0, 0.0, prec=2)["info"] == 10
    nodiag = (np.array([0, 1, 2], np.int32), np.array([1, 0], np.int32), np.array([1.0, 1.0]))
    assert O.ilu0(nodiag)[0] == 10


def test_cg_ilu0_dense_solve_and_fewer_iterations():
    """BJAC-preconditioned shifted CG solves (eta A + zeta I) x = b, with the
    factors taken from the unshifted A, and needs fewer iterations than DIAG
    on an anisotropic operator."""
    A = matgen.anisotropic_2d(12, seed=4)
    n = 144
    Ad = matgen.to_scipy(A).toarray()
    b = matgen.uniform_pm1(3, n)
    eta, zeta = 1.0, 0.1
    ex = np.linalg.solve(eta * Ad + zeta * np.eye(n), b)
    rb = O.cg_shifted(A, b, eta, zeta, eps=1e-12, itmax=2000, istop=2, prec=2)
    rd = O.cg_shifted(A, b, eta, zeta, eps=1e-12, itmax=2000, istop=2, prec=1)
    assert rb["info"] == 0 and np.linalg.norm(rb["x"] - ex) <= 1e-10 * np.linalg.norm(ex)
    assert rb["iter"] < rd["iter"] / 2


# ---- multi-shift CG, quadrature --------------------------------------------
def test_mscg_dense_solve():
    A = matgen.laplacian((8, 8))
    Ad = matgen.to_scipy(A).toarray()
    b = matgen.uniform_pm1(9, 64)
    shifts = np.array([0.0, 0.5, 2.0 + 1.0j, -0.1 + 3.0j, 5.0 - 2.0j])
    r = O.mscg(A, b, shifts, eps=1e-12, itmax=500)
    for s, xs in zip(shifts, r["X"]):
        ex = np.linalg.solve(Ad + s * np.eye(64), b)
        assert np.linalg.norm(xs - ex) < 1e-9 * np.linalg.norm(ex)


def test_talbot_scalar_accuracy():
    s, w = O.talbot(16, 1.0)
    xs = np.linspace(0, 12, 400)
    approx = np.array([np.real(np.sum(w / (s + x))) for x in xs])
    assert np.max(np.abs(approx - np.exp(-xs))) < 2e-9
    s, w = O.talbot(24, 1.0)
    approx = np.array([np.real(np.sum(w / (s + x))) for x in xs])
    assert np.max(np.abs(approx - np.exp(-xs))) < 1e-12


def test_talbot_quad_apply_expm():
    g = load_golden("quad_small.npz")
    A = (g["rowptr"], g["colind"], g["val"])
    s, w = O.talbot(24, 1.0)
    r = O.quad_apply(A, s, w, 1.0, 0, g["x"], eps=1e-13, itmax=500)
    assert np.linalg.norm(r["y"] - g["y_expm"]) < 1e-10 * np.linalg.norm(g["y_expm"])


def test_hht_method1_sqrt():
    g = load_golden("quad_small.npz")
    A = (g["rowptr"], g["colind"], g["val"])
    ev = np.linalg.eigvalsh(matgen.to_scipy(A).toarray())
    xi, c, eta, sign = O.hhtmethod1(np.sqrt, 24, [ev[0], ev[-1]])
    assert sign == -1.0
    r = O.quad_apply(A, sign * xi, c, eta * sign, 1, g["x"], eps=1e-14, itmax=800)
    assert np.linalg.norm(r["y"] - g["y_sqrt"]) < 1e-9 * np.linalg.norm(g["y_sqrt"])


def test_hht_quadraturetest_poles():
    # examples/quadrature/quadraturetest.F90:60-67: N = 10, [0.01, 4], f = sqrt(1+z)
    xi, c, eta, sign = O.hhtmethod1(lambda z: np.sqrt(1 + z), 10, [0.01, 4.0])
    # method 1 places the N poles on the upper half of the contour (the
    # imaginary part in the assembly accounts for the conjugate half)
    assert np.all(xi.imag > 0) and eta < 0 and sign == -1.0
    assert abs(eta - (-4 * O.ellipkkp(-np.log((np.sqrt(400) - 1) / (np.sqrt(400) + 1)) / np.pi)[0] * 0.2 /
                      (((np.sqrt(400) - 1) / (np.sqrt(400) + 1)) * np.pi * 10))) < 1e-14
    # the rule integrates sqrt(1+z) at scalar points of the interval (N = 10: ~1e-4)
    for lam in (0.02, 0.5, 3.9):
        v = eta * lam * np.imag(np.sum(c / (xi - lam)))
        assert abs(v - np.sqrt(1 + lam)) < 1e-4
    # and converges as N grows
    xi, c, eta, sign = O.hhtmethod1(lambda z: np.sqrt(1 + z), 40, [0.01, 4.0])
    for lam in (0.02, 0.5, 3.9):
        assert abs(eta * lam * np.imag(np.sum(c / (xi - lam))) - np.sqrt(1 + lam)) < 1e-12


def test_dense_inverse_is_inverse():
    rng = np.random.default_rng(3)
    for k in (1, 2, 7, 31):
        H = rng.normal(size=(k, k)) + k * np.eye(k)
        Hi = O.dense_inverse(H)
        assert np.allclose(H @ Hi, np.eye(k), rtol=0, atol=1e-12)
        assert np.allclose(Hi, np.linalg.inv(H), rtol=1e-12, atol=1e-13)
    # a permutation needs pivoting
    P = np.eye(4)[[2, 0, 3, 1]]
    assert np.array_equal(O.dense_inverse(P), P.T)


@pytest.mark.parametrize("tau", [0.0, -5.0])
def test_sai_arnoldi_closed_form(tau):
    """Shift-and-invert Arnoldi on the stiff A = -20 L (negative definite, so
    (A + tau I) is definite for tau <= 0 and the inner CG applies): exp(A) x
    against the DST closed form (inner solves to the outer eps 1e-10 bound
    the accuracy near 1e-8), in far fewer steps than polynomial Arnoldi."""
    dims = (20, 20)
    A = matgen.laplacian(dims, scale=-20.0)
    x = matgen.uniform_pm1(17, 400)
    r = O.sai_arnoldi(A, x, 1e-10, 150, tau=tau, istop=1, in_itmax=4000, in_istop=2)
    assert r["info"] == 0
    yex = matgen.dst_expm_laplacian(dims, x, 20.0)
    assert np.linalg.norm(r["y"] - yex) <= 2e-8 * np.linalg.norm(yex)
    ra = O.arnoldi(A, x, 1e-10, 150, 1)
    assert r["iter"] < 0.5 * ra["iter"]
    assert abs(r["res"][0] - np.linalg.norm(x)) <= 1e-12 * np.linalg.norm(x)


# ---- USERF/SYM f(H) step (serial/psfun_d_serial_mod.F90:420-485) ----------
@pytest.mark.parametrize("k", [1, 2, 5, 33])
def test_sym_eig_matches_lapack(k):
    """Eigenvalues of the symmetric matrix of H's upper triangle against
    numpy's LAPACK eigh; Z orthonormal; Z diag(w) Z^T reproduces it."""
    rng = np.random.default_rng(100 + k)
    H = rng.normal(size=(k, k))
    S = np.triu(H) + np.triu(H, 1).T
    w, Z = O.sym_eig(H)
    we = np.linalg.eigh(S)[0]
    scale = max(1.0, np.abs(we).max())
    assert np.abs(np.sort(w) - we).max() <= 1e-14 * k * scale
    assert np.linalg.norm(Z.T @ Z - np.eye(k)) <= 1e-13 * k
    assert np.linalg.norm(Z @ np.diag(w) @ Z.T - S) <= 1e-14 * k * np.linalg.norm(S)


def test_userf_sym_apply_closed_forms():
    """f = exp against scipy expm, f = sqrt against scipy sqrtm (SPD), f = 1/z
    against a dense solve, each of the upper-triangle matrix with a scaling."""
    rng = np.random.default_rng(7)
    k = 24
    H = rng.normal(size=(k, k))
    S = np.triu(H) + np.triu(H, 1).T
    x = rng.normal(size=k)
    y = O.userf_sym_apply(H, -0.4, np.exp, x)
    assert np.linalg.norm(y - sla.expm(-0.4 * S) @ x) <= 1e-13 * np.linalg.norm(y)
    B = rng.normal(size=(k, k))
    P = B @ B.T + k * np.eye(k)
    Pu = np.triu(P) + 3.0 * np.tril(P, -1)     # lower triangle must be ignored
    y = O.userf_sym_apply(Pu, 2.0, np.sqrt, x)
    assert np.linalg.norm(y - np.real(sla.sqrtm(2.0 * P)) @ x) <= 1e-12 * np.linalg.norm(y)
    y = O.userf_sym_apply(P, 1.0, lambda z: 1.0 / z, x)
    assert np.linalg.norm(y - np.linalg.solve(P, x)) <= 1e-12 * np.linalg.norm(y)


def test_lanczos_userf_exp_equals_exp():
    """USERF with f = exp reproduces the EXP Lanczos: same iterations, same
    estimates, same result to rounding."""
    A = matgen.laplacian((12, 12), scale=-1.0)
    x = matgen.uniform_pm1(3, 144)
    a = O.lanczos(A, x, 1e-10, 60, 1)
    b = O.lanczos(A, x, 1e-10, 60, 1, f=np.exp)
    assert a["iter"] == b["iter"]
    assert np.allclose(a["res"], b["res"], rtol=1e-8, atol=1e-13)
    assert np.linalg.norm(a["y"] - b["y"]) <= 1e-12 * np.linalg.norm(a["y"])


def test_lanczos_userf_sqrt_against_sqrtm():
    """f = sqrt of the SPD Laplacian (opscale +1) against scipy sqrtm."""
    A = matgen.laplacian((9, 9), scale=1.0)
    x = matgen.uniform_pm1(4, 81)
    r = O.lanczos(A, x, 1e-12, 81, 1, f=np.sqrt)
    Ad = matgen.to_scipy(A).toarray()
    ref = np.real(sla.sqrtm(Ad)) @ x
    assert np.linalg.norm(r["y"] - ref) <= 1e-9 * np.linalg.norm(ref)


def test_arnoldi_and_sai_userf_exp_equal_exp():
    """On a symmetric A, H_k is symmetric up to rounding, so USERF/SYM with
    f = exp reproduces the EXP Arnoldi and SAI projections."""
    A = matgen.laplacian((10, 10), scale=-1.0)
    x = matgen.uniform_pm1(5, 100)
    a = O.arnoldi(A, x, 1e-10, 60, 1)
    b = O.arnoldi(A, x, 1e-10, 60, 1, f=np.exp)
    assert a["iter"] == b["iter"]
    assert np.linalg.norm(a["y"] - b["y"]) <= 1e-11 * np.linalg.norm(a["y"])
    A = matgen.laplacian((10, 10), scale=-10.0)
    a = O.sai_arnoldi(A, x, 1e-10, 60, tau=0.0, istop=1, in_itmax=2000)
    b = O.sai_arnoldi(A, x, 1e-10, 60, tau=0.0, istop=1, in_itmax=2000, f=np.exp)
    assert abs(a["iter"] - b["iter"]) <= 1
    assert np.linalg.norm(a["y"] - b["y"]) <= 1e-9 * np.linalg.norm(a["y"])
    assert np.linalg.norm(b["y"] - matgen.dst_expm_laplacian((10, 10), x, 10.0)) <= 1e-8 * np.linalg.norm(b["y"])


# ---- preconditioned COCG: the per-shift fallback of the multi-shift solve ---
def _aniso(m, seed):
    A = matgen.anisotropic_2d(m, seed=seed)
    return A, matgen.to_scipy(A).toarray()


@pytest.mark.parametrize("prec", [0, 1, 2])
@pytest.mark.parametrize("sigma", [0.3 + 0.2j, -0.5 + 0.01j, 4.0 - 3.0j])
def test_pcocg_dense_solve(prec, sigma):
    """Each preconditioner solves (A + sigma I) x = b, also for an
    indefinite shift (-0.5 lies inside the spectrum [0.35, 221]); resid is
    the true relative residual."""
    A, Ad = _aniso(12, 4)
    n = 144
    b = matgen.uniform_pm1(2, n)
    K = Ad + sigma * np.eye(n)
    r = O.pcocg(A, b, sigma, prec=prec, eps=1e-12, itmax=3000)
    ex = np.linalg.solve(K, b)
    assert r["info"] == 0 and r["resid"] <= 1e-12 and 0 < r["iter"] < 3000
    assert np.linalg.norm(r["x"] - ex) <= 1e-9 * np.linalg.norm(ex)
    true = np.linalg.norm(b - K @ r["x"]) / np.linalg.norm(b)
    assert abs(true - r["resid"]) <= 1e-13


def test_pcocg_without_preconditioner_is_the_multishift_iterate():
    """With M = I and x0 = 0, COCG on one shift and the multi-shift CG build
    the same Krylov space under the same Galerkin condition, so after k
    steps their iterates agree: a wrong sign or index in either recurrence
    breaks it."""
    A = matgen.laplacian((9, 8))
    b = matgen.uniform_pm1(6, 72)
    shifts = [0.5 + 1.0j, 2.0 - 0.3j, 0.05 + 0.0j]
    for k in (1, 4, 9):
        ms = O.mscg(A, b, shifts, eps=1e-30, itmax=k)
        for s, xs in zip(shifts, ms["X"]):
            r = O.pcocg(A, b, s, prec=0, eps=1e-30, itmax=k)
            assert r["iter"] == k
            assert np.linalg.norm(r["x"] - xs) <= 1e-11 * np.linalg.norm(xs), (k, s)


def test_pcocg_warm_start_and_edge_cases():
    A, Ad = _aniso(8, 1)
    n = 64
    b = matgen.uniform_pm1(3, n)
    s = 0.7 + 0.4j
    ex = np.linalg.solve(Ad + s * np.eye(n), b)
    r = O.pcocg(A, b, s, x0=ex, prec=2, eps=1e-10)
    assert r["iter"] == 0 and r["resid"] <= 1e-10 and np.array_equal(r["x"], ex)
    half = 0.5 * ex
    r = O.pcocg(A, b, s, x0=half, prec=1, eps=1e-12)
    assert r["iter"] > 0 and np.linalg.norm(r["x"] - ex) <= 1e-9 * np.linalg.norm(ex)
    r = O.pcocg(A, np.zeros(n), s, x0=half, prec=1)
    assert r["iter"] == 0 and r["resid"] == 0.0 and np.all(r["x"] == 0)
    # a zero entry of diag(A) + sigma is replaced by 1 (sigma = -diag cancels it)
    D = (np.arange(3, dtype=np.int32), np.arange(2, dtype=np.int32), np.array([2.0, 5.0]))
    r = O.pcocg(D, np.array([1.0, 1.0]), -2.0, prec=1, eps=1e-14, itmax=10)
    assert r["info"] == 0 and np.all(np.isfinite(r["x"]))


def test_bjac_is_symmetric_for_a_symmetric_matrix():
    """COCG needs a (complex) symmetric M: for a symmetric A the ILU(0)
    factors satisfy U = D L^T, so M = L U = L D L^T."""
    A, Ad = _aniso(9, 2)
    rc, rp, ci, lu = O.ilu0(A)
    L, U = _ilu_factors(rp, ci, lu)
    M = L @ U
    assert rc == 0 and np.max(np.abs(M - M.T)) <= 1e-13 * np.max(np.abs(M))


def _stalled_case():
    """sqrt(A) x by HHT method 1 (20 poles) on an anisotropic operator with
    kappa ~ 1.7e3 at eps 1e-12: 60 multi-shift iterations leave the poles
    nearest the spectrum unconverged."""
    A, Ad = _aniso(20, 6)
    ev = np.linalg.eigvalsh(Ad)
    xi, c, eta, sign = O.hhtmethod1(np.sqrt, 20, [ev[0], ev[-1]])
    return A, Ad, sign * xi, c, eta * sign


@pytest.mark.parametrize("prec,itmax", [(1, 150), (2, 60)])
def test_mscg_prec_finishes_the_stalled_shifts(prec, itmax):
    A, Ad, shifts, _, _ = _stalled_case()
    n = len(Ad)
    b = matgen.uniform_pm1(11, n)
    plain = O.mscg(A, b, shifts, eps=1e-12, itmax=itmax)
    stalled = plain["resid"] > 1e-12
    assert stalled.sum() >= 5 and (~stalled).sum() >= 5
    r = O.mscg_prec(A, b, shifts, prec, eps=1e-12, itmax=itmax)
    assert r["info"] == 0 and np.all(r["resid"] <= 1e-12)
    # without a preconditioner the same budget leaves shifts unfinished
    assert np.any(O.mscg_prec(A, b, shifts, 0, eps=1e-12, itmax=itmax)["resid"] > 1e-12)
    # the shifts the multi-shift solve finished are its own result, unchanged
    assert np.array_equal(r["X"][~stalled], plain["X"][~stalled])
    assert np.array_equal(r["iters"][~stalled], plain["iters"][~stalled])
    assert np.all(r["iters"][stalled] > itmax)
    for s, xs in zip(shifts, r["X"]):
        ex = np.linalg.solve(Ad + s * np.eye(n), b)
        assert np.linalg.norm(xs - ex) <= 1e-9 * np.linalg.norm(ex)


def test_quad_apply_prec_sqrt_against_sqrtm():
    A, Ad, shifts, c, gamma = _stalled_case()
    x = matgen.uniform_pm1(11, len(Ad))
    ysq = np.real(sla.sqrtm(Ad)) @ x
    plain = O.quad_apply(A, shifts, c, gamma, 1, x, eps=1e-12, itmax=60)
    assert np.linalg.norm(plain["y"] - ysq) > 1e-5 * np.linalg.norm(ysq)
    for prec, itmax in ((1, 150), (2, 60)):
        r = O.quad_apply(A, shifts, c, gamma, 1, x, eps=1e-12, itmax=itmax, prec=prec)
        # the 20-pole rule's own error on this spectrum is 5.4e-9
        assert r["info"] == 0 and np.linalg.norm(r["y"] - ysq) <= 1e-8 * np.linalg.norm(ysq)


# ---- GENPADE with fun%padedegree: DGPADM's Pade approximant ----------------
def _pade_scalar(x, p):
    from math import factorial
    c = [factorial(2 * p - j) * factorial(p) / (factorial(2 * p) * factorial(j) * factorial(p - j))
         for j in range(p + 1)]
    return sum(c[j] * x ** j for j in range(p + 1)) / sum(c[j] * (-x) ** j for j in range(p + 1))


def _pade_ns(hnorm):
    return max(0, int(np.log(hnorm) / np.log(2.0)) + 2)    # int() truncates toward zero, as Fortran INT


@pytest.mark.parametrize("p", [2, 3, 4, 6, 9])
@pytest.mark.parametrize("h", [0.3, -0.7, 5.0, -40.0, 1e-3])
def test_pade_expm_scalar_closed_form(p, h):
    """1x1: r_p(h / 2^ns)^(2^ns) with the Pade coefficients written from
    factorials, independent of the even/odd Horner split."""
    ref = _pade_scalar(h / 2.0 ** _pade_ns(abs(h)), p) ** (2 ** _pade_ns(abs(h)))
    rc, E = O.pade_expm([[h]], 1.0, p)
    assert rc == 0 and abs(E[0, 0] - ref) <= 1e-13 * abs(ref)


@pytest.mark.parametrize("p", [2, 5])
def test_pade_expm_diagonal_shares_one_scaling(p):
    d = np.array([0.5, -3.0, 1.25, 7.0])
    t = -0.5
    ns = _pade_ns(abs(t) * 7.0)
    rc, E = O.pade_expm(np.diag(d), t, p)
    ref = [_pade_scalar(t * v / 2.0 ** ns, p) ** (2 ** ns) for v in d]
    assert rc == 0 and np.allclose(np.diag(E), ref, rtol=1e-13, atol=0) and np.all(E - np.diag(np.diag(E)) == 0)


@pytest.mark.parametrize("p", [6, 8, 13])
def test_pade_expm_matches_scipy_at_full_degree(p):
    rng = np.random.default_rng(p)
    for k in (6, 20):
        H = rng.standard_normal((k, k)) / np.sqrt(k) * 2.0
        rc, E = O.pade_expm(H, 1.3, p)
        ex = sla.expm(1.3 * H)
        assert rc == 0 and np.linalg.norm(E - ex) <= 1e-13 * np.linalg.norm(ex) * k


def test_pade_expm_low_degree_carries_its_truncation_and_zero_is_identity():
    rng = np.random.default_rng(3)
    H = rng.standard_normal((10, 10))
    rc, E2 = O.pade_expm(H, 1.0, 2)
    ex = sla.expm(H)
    assert rc == 0 and np.linalg.norm(E2 - ex) > 1e-7 * np.linalg.norm(ex)   # 5.3e-7 here
    assert O.pade_expm(np.zeros((3, 3)), 1.0, 4)[1].tolist() == np.eye(3).tolist()
    assert O.pade_expm(H, 1.0, 1)[0] == 2
