"""GPU parity of the PSBLAS primitives, the f(T) step and the Lanczos f(A)x
against the CPU oracle, the closed-form DST solution and the golden fixtures.

Bars: integer/index work exact; FP64 primitives to a few ulps; Lanczos
y within relative 2-norm 1e-10 of the oracle (the reference's polynomial
path tolerance)."""
import numpy as np
import pytest
import torch

from conftest import load_golden
from oracle import oracle as O
from psfun_amd import matgen, psb
from psfun_amd.krylov import psfun_d_krylov
from psfun_amd.serial import psfun_d_serial

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, np.float64), device="cuda:0")


def _vec(a):
    return psb.psb_d_vect_type(_dev(a))


def _ragged(n, seed):
    """CSR with empty rows, long rows and a dense row (ragged lengths)."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 12, n)
    lens[::17] = 0
    lens[5 % n] = min(n, 150)
    rows, cols = [], []
    for i, L in enumerate(lens):
        c = np.sort(rng.choice(n, size=min(L, n), replace=False))
        rows += [i] * len(c)
        cols += list(c)
    rows, cols = np.array(rows, np.int64), np.array(cols, np.int64)
    vals = rng.normal(size=len(rows))
    return matgen._coo_to_csr(n, rows, cols, vals)


@pytest.mark.parametrize("lanes", [0, 1, 2, 4, 8, 16, 32, 64])
def test_spmv_ragged_all_lane_widths(gpu_ctx, lanes):
    A = _ragged(3001, 1)
    M = psb.psb_dspmat_type.from_csr(A)
    if lanes:
        M.set_lanes(lanes)
    x = matgen.uniform_pm1(2, 3001)
    y0 = matgen.uniform_pm1(3, 3001)
    desc = psb.psb_desc_type(3001)
    for alpha, beta in ((1.0, 0.0), (-0.5, 2.0)):
        y = _vec(y0)
        assert psb.psb_spmm(alpha, M, _vec(x), beta, y, desc) == 0
        ref = O.spmv(A, x, alpha, beta, y0)
        assert np.max(np.abs(y.get_vect() - ref)) <= 1e-13 * (np.max(np.abs(ref)) + 1)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 3001])
def test_spmv_sliced_ell_matches_oracle(gpu_ctx, n):
    """Sliced ELL (64-row slices): partial last slice, empty rows, one long
    row (heavy padding), and a value refresh through set_values."""
    A = _ragged(n, 11)
    M = psb.psb_dspmat_type.from_csr(A)
    M.set_format("SELL")
    fmt, entries = M.get_format()
    nslice = (n + 63) // 64
    lens = np.diff(A[0])
    widths = [lens[64 * s:64 * s + 64].max() for s in range(nslice)]
    assert fmt == "SELL" and entries == 64 * sum(widths)
    x = matgen.uniform_pm1(12, n)
    y0 = matgen.uniform_pm1(13, n)
    desc = psb.psb_desc_type(n)
    for alpha, beta in ((1.0, 0.0), (-0.5, 2.0)):
        y = _vec(y0)
        assert psb.psb_spmm(alpha, M, _vec(x), beta, y, desc) == 0
        ref = O.spmv(A, x, alpha, beta, y0)
        assert np.max(np.abs(y.get_vect() - ref)) <= 1e-13 * (np.max(np.abs(ref)) + 1)
    v2 = matgen.normal(14, len(A[2]))
    M.set_values(_dev(v2))
    y = _vec(np.zeros(n))
    psb.psb_spmm(1.0, M, _vec(x), 0.0, y, desc)
    ref = O.spmv((A[0], A[1], v2), x)
    assert np.max(np.abs(y.get_vect() - ref)) <= 1e-13 * (np.max(np.abs(ref)) + 1)


def _banded(n, half, drop, seed, dup=False):
    """Band matrix with offsets -half..half, a fraction `drop` of the entries
    removed at random (so slices hold diagonals some rows lack, and the first
    and last rows' slices hold offsets that run off the matrix); with dup, one
    entry per 97 rows is stored twice."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for i in range(n):
        for d in range(-half, half + 1):
            c = i + d
            if 0 <= c < n and rng.random() >= drop:
                rows.append(i)
                cols.append(c)
                if dup and i % 97 == 0 and d == 0:
                    rows.append(i)
                    cols.append(c)
    rows, cols = np.array(rows, np.int64), np.array(cols, np.int64)
    return rows, cols, rng.normal(size=len(rows))


def _csr_keep_duplicates(n, rows, cols, vals):
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, rows + 1, 1)
    return np.cumsum(rp).astype(np.int32), cols.astype(np.int32), vals


@pytest.mark.parametrize("n", [1, 63, 64, 65, 3001])
def test_spmv_diagonal_slices_match_oracle(gpu_ctx, n):
    """Diagonal slices (hacked DIA): diagonals absent from some rows, offsets
    running off both ends of the matrix (clamped reads of the zero padding), a
    partial last slice, repeated entries summed, and a value refresh."""
    rows, cols, vals = _banded(n, 3, 0.1, 21, dup=True)
    A = _csr_keep_duplicates(n, rows, cols, vals)
    M = psb.psb_dspmat_type.from_csr(A)
    M.set_format("HDIA")
    assert M.get_format()[0] == "HDIA"
    nslice = (n + 63) // 64
    expect = sum(len(np.unique(cols[(rows >= 64 * s) & (rows < 64 * s + 64)] -
                               rows[(rows >= 64 * s) & (rows < 64 * s + 64)])) for s in range(nslice))
    assert M.hdia_diagonals() == expect
    x = matgen.uniform_pm1(22, n)
    y0 = matgen.uniform_pm1(23, n)
    desc = psb.psb_desc_type(n)
    for alpha, beta in ((1.0, 0.0), (-0.5, 2.0)):
        y = _vec(y0)
        assert psb.psb_spmm(alpha, M, _vec(x), beta, y, desc) == 0
        ref = O.spmv(A, x, alpha, beta, y0)
        assert np.max(np.abs(y.get_vect() - ref)) <= 1e-13 * (np.max(np.abs(ref)) + 1)
    v2 = matgen.normal(24, len(A[2]))
    M.set_values(_dev(v2))
    y = _vec(np.zeros(n))
    psb.psb_spmm(1.0, M, _vec(x), 0.0, y, desc)
    ref = O.spmv((A[0], A[1], v2), x)
    assert np.max(np.abs(y.get_vect() - ref)) <= 1e-13 * (np.max(np.abs(ref)) + 1)


def test_diagonal_slices_refused_for_scattered_pattern(gpu_ctx):
    """Random columns give every slice hundreds of diagonals: the format is
    refused (PSFUN_ERR_UNSUPPORTED) and the previous choice stays."""
    from psfun_amd import _lib
    M = psb.psb_dspmat_type.from_csr(_ragged(3001, 1))
    M.set_format("SELL")
    with pytest.raises(_lib.PsfunError) as e:
        M.set_format("HDIA")
    assert e.value.status == _lib.PSFUN_ERR_UNSUPPORTED
    assert M.get_format()[0] == "SELL" and M.hdia_diagonals() == 0


def test_lanczos_diagonal_slices_match_sliced_ell(gpu_ctx):
    """The Lanczos step on diagonal slices adds each row's terms in the same
    order as the sliced-ELL step: the same iterations and y to rounding."""
    A = matgen.laplacian((20, 20, 20), scale=-1.0)
    x = matgen.uniform_pm1(25, 8000)
    r = {f: _lanczos_gpu(A, x, 1e-10, 50, fmt=f) for f in ("SELL", "HDIA")}
    assert r["HDIA"][0] == 0 and r["HDIA"][1] == r["SELL"][1]
    assert np.linalg.norm(r["HDIA"][4] - r["SELL"][4]) <= 1e-14 * np.linalg.norm(r["SELL"][4])


def test_format_auto_choice(gpu_ctx):
    M = psb.psb_dspmat_type.from_csr(matgen.laplacian((40, 40, 3)))
    assert M.get_format()[0] == "SELL"        # uniform rows: padding ~0
    M = psb.psb_dspmat_type.from_csr(_ragged(3001, 1))
    assert M.get_format()[0] == "CSR"         # one 150-entry row pads its slice
    M.set_format("SELL")
    assert M.get_format()[0] == "SELL"


def test_lanczos_sell_index_widths(gpu_ctx):
    """The Lanczos sliced-ELL step reads 16-bit column offsets when every
    column is within 32767 of its row, 32-bit columns otherwise; both match
    the oracle and the CSR path."""
    A = matgen.laplacian((20, 20, 20), scale=-1.0)               # offsets up to 400
    M = psb.psb_dspmat_type.from_csr(A)
    M.set_format("SELL")
    assert M.sell_index_bits() == 16
    # a path graph of 40 000 nodes closed by one long edge (offset 39 999)
    n = 40000
    rows = np.concatenate([np.arange(n - 1), np.arange(1, n), np.arange(n), [0, n - 1]])
    cols = np.concatenate([np.arange(1, n), np.arange(n - 1), np.arange(n), [n - 1, 0]])
    vals = np.concatenate([np.ones(2 * n - 2), -2.5 * np.ones(n), [0.5, 0.5]])
    B = matgen._coo_to_csr(n, rows, cols, vals)
    MB = psb.psb_dspmat_type.from_csr(B)
    MB.set_format("SELL")
    assert MB.sell_index_bits() == 32
    for mat, nn in ((A, 8000), (B, n)):
        x = matgen.uniform_pm1(5, nn)
        ref = O.lanczos(mat, x, 1e-10, 50, 1)
        ys = {}
        for fmt in ("SELL", "CSR"):
            info, it, err, res, y = _lanczos_gpu(mat, x, 1e-10, 50, fmt=fmt)
            assert info == 0 and abs(it - ref["iter"]) <= 1
            assert np.linalg.norm(y - ref["y"]) <= 1e-10 * np.linalg.norm(ref["y"])
            ys[fmt] = y
        assert np.linalg.norm(ys["SELL"] - ys["CSR"]) <= 1e-12 * np.linalg.norm(ys["CSR"])


def test_spmv_nonfinite_x_same_in_every_format(gpu_ctx):
    """One Inf and one NaN entry of x: in every storage format exactly the
    rows whose stored entries meet them are non-finite (the padding of
    sliced ELL and diagonal slices never multiplies a real x entry); the
    Lanczos SpMV steps read the same row sums."""
    dims = (20, 13)
    A = matgen.laplacian(dims)
    n = 260
    x = matgen.uniform_pm1(12, n)
    x[57], x[200] = np.inf, np.nan
    Asp = matgen.to_scipy(A)
    hit = np.asarray(Asp[:, [57, 200]].getnnz(axis=1) > 0).ravel()
    ref = O.spmv(A, x)
    assert np.array_equal(~np.isfinite(ref), hit)
    desc = psb.psb_desc_type(n)
    for fmt in ("CSR", "SELL", "HDIA"):
        M = psb.psb_dspmat_type.from_csr(A)
        M.set_format(fmt)
        y = _vec(np.zeros(n))
        assert psb.psb_spmm(1.0, M, _vec(x), 0.0, y, desc) == 0
        got = y.get_vect()
        assert np.array_equal(~np.isfinite(got), hit), fmt
        assert np.array_equal(np.isnan(got), np.isnan(ref)), fmt
        assert np.max(np.abs(got[~hit] - ref[~hit])) <= 1e-13, fmt


def test_spmv_one_based_input(gpu_ctx):
    g = load_golden("laplacian50.npz")
    M = psb.psb_dspmat_type(g["rowptr"], g["colind"], g["val"], index_base=1)
    x = np.arange(50, dtype=np.float64)
    y = _vec(np.zeros(50))
    psb.psb_spmm(1.0, M, _vec(x), 0.0, y, psb.psb_desc_type(50))
    A0 = (g["rowptr"] - 1, g["colind"] - 1, g["val"])
    assert np.array_equal(y.get_vect(), O.spmv(A0, x))


def test_blas1(gpu_ctx):
    n = 1_000_003
    a = matgen.uniform_pm1(4, n)
    b = matgen.normal(5, n)
    desc = psb.psb_desc_type(n)
    va, vb = _vec(a), _vec(b)
    d, info = psb.psb_gedot(va, vb, desc)
    assert info == 0 and abs(d - a @ b) <= 1e-12 * np.abs(a * b).sum()
    r, _ = psb.psb_genrm2(va, desc)
    assert abs(r - np.linalg.norm(a)) <= 1e-13 * np.linalg.norm(a)
    m, _ = psb.psb_geamax(vb, desc)
    assert m == np.max(np.abs(b))
    assert psb.psb_geaxpby(2.0, va, -3.0, vb, desc) == 0
    assert np.max(np.abs(vb.get_vect() - (2 * a - 3 * b))) < 1e-14 * 10
    assert psb.psb_scal(0.5, va, desc) == 0
    assert np.array_equal(va.get_vect(), 0.5 * a)
    # empty vectors
    e = psb.psb_d_vect_type(torch.zeros(0, dtype=torch.float64, device="cuda:0"))
    assert psb.psb_gedot(e, e, psb.psb_desc_type(0)) == (0.0, 0)


def test_serial_fT_step_against_high_precision(gpu_ctx):
    g = load_golden("expm_small.npz")
    fun = psfun_d_serial()
    fun.set("VARIANT", "CHBSYM")
    for k in (1, 2, 5, 17, 40, 64, 65, 130):
        for s in (1.0, -0.5):
            d, e, y = g[f"tri_{k}_{s}_d"], g[f"tri_{k}_{s}_e"], g[f"tri_{k}_{s}_y"]
            T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
            fun.set("SCALING", s)
            out = np.zeros(k)
            e1 = np.zeros(k)
            e1[0] = 1
            assert fun.apply(T, out, e1) == 0
            assert np.linalg.norm(out - y) <= 1e-12 * np.linalg.norm(y), (k, s)
            # the Chebyshev recurrence keeps the trailing components to relative
            # accuracy down to the series truncation (terms below 1e-26 of the first)
            assert abs(out[-1] - y[-1]) <= 1e-9 * abs(y[-1]) + 1e-24 * np.linalg.norm(y), \
                (k, s, out[-1], y[-1])


def test_serial_rejects_default_variant(gpu_ctx):
    fun = psfun_d_serial()           # variant 'EXPOKIT' is not a valid case
    out = np.zeros(3)
    assert fun.apply(np.eye(3), out, np.ones(3)) == psb.psb_err_from_subroutine_
    fun.set("VARIANT", "chbsym")     # values are case-sensitive
    assert fun.apply(np.eye(3), out, np.ones(3)) == psb.psb_err_from_subroutine_


def _lanczos_gpu(A, x, eps, itmax, istop=1, fmt="AUTO"):
    n = len(x)
    M = psb.psb_dspmat_type.from_csr(A)
    M.set_format(fmt)
    desc = psb.psb_desc_type(n)
    fun = psfun_d_serial()
    fun.set("VARIANT", "CHBSYM")
    km = psfun_d_krylov()
    km.set("KNAME", "LANCZOS")
    y = _vec(np.zeros(n))
    info, it, err, res = km.apply(fun, M, desc, y, _vec(x), eps, itmax=itmax, istop=istop)
    return info, it, err, res, y.get_vect()


@pytest.mark.parametrize("dims,itmax,fmt", [((100, 100), 30, "AUTO"), ((128, 128, 128), 50, "HDIA")])
def test_lanczos_stop_index_matches_reference_cram(gpu_ctx, dims, itmax, fmt):
    """The reference's CHBSYM computes f(T) by EXPOKIT's CRAM(14,14) (the
    oracle's exp_cram mode); the device's exact f(T) stops at the same index
    on the CPU-runnable config (100^2) and the plain GPU config (128^3, on
    the diagonal slices the bench times), with the same final estimate to
    CRAM's error and y within that error of the reference's."""
    print(f"[cram] dims={dims} fmt={fmt}", flush=True)
    A = matgen.laplacian(dims, scale=-1.0)
    x = matgen.uniform_pm1(1, int(np.prod(dims)))
    info, it, err, res, y = _lanczos_gpu(A, x, 1e-10, itmax, fmt=fmt)
    with O.exp_cram():
        ref = O.lanczos(A, x, 1e-10, itmax, 1)
    assert info == 0 and it == ref["iter"]
    assert err < 1e-10 and ref["err"] < 1e-10
    assert np.allclose(res[1:it - 2], ref["res"][1:it - 2], rtol=1e-2, atol=0)
    assert np.linalg.norm(y - ref["y"]) <= 1e-11 * np.linalg.norm(ref["y"])


@pytest.mark.parametrize("fmt", ["CSR", "SELL", "HDIA", "CSRS"])
@pytest.mark.parametrize("dims,itmax", [((100, 100), 30), ((12, 10, 9), 50), ((64, 64, 64), 50)])
def test_lanczos_parity_closed_form(gpu_ctx, dims, itmax, fmt):
    A = matgen.laplacian(dims, scale=-1.0)
    x = matgen.uniform_pm1(1, int(np.prod(dims)))
    info, it, err, res, y = _lanczos_gpu(A, x, 1e-10, itmax, fmt=fmt)
    assert info == 0
    ref = O.lanczos(A, x, 1e-10, itmax, 1)
    rel = np.linalg.norm(y - ref["y"]) / np.linalg.norm(ref["y"])
    assert rel <= 1e-10, rel
    yex = matgen.dst_expm_laplacian(dims, x, 1.0)
    assert np.linalg.norm(y - yex) / np.linalg.norm(yex) <= 1e-10
    assert abs(it - ref["iter"]) <= 1
    assert abs(res[0] - np.linalg.norm(x)) <= 1e-12 * np.linalg.norm(x)
    assert len(res) == it + 1 and err < 1e-10


@pytest.mark.parametrize("fmt", ["CSR", "SELL", "HDIA", "CSRS"])
def test_lanczos_long_run_lds_fT(gpu_ctx, fmt):
    """exp(-30 L) on 16^3 takes ~90 iterations: the loop runs with itmax > 65,
    so the control block's f(T) step is the LDS routine for orders past 64."""
    dims = (16, 16, 16)
    A = matgen.laplacian(dims, scale=-30.0)
    x = matgen.uniform_pm1(8, 4096)
    info, it, err, res, y = _lanczos_gpu(A, x, 1e-10, 200, fmt=fmt)
    assert info == 0 and it > 66
    ref = O.lanczos(A, x, 1e-10, 200, 1)
    assert abs(it - ref["iter"]) <= 1
    assert np.linalg.norm(y - ref["y"]) <= 1e-10 * np.linalg.norm(ref["y"])
    assert np.linalg.norm(y - matgen.dst_expm_laplacian(dims, x, 30.0)) <= 1e-9 * np.linalg.norm(y)


def test_lanczos_reference_example(gpu_ctx):
    g = load_golden("laplacian50.npz")
    A = (g["rowptr"] - 1, g["colind"] - 1, g["val"])
    info, it, err, res, y = _lanczos_gpu(A, np.ones(50), 1e-6, 100)
    ref = O.lanczos(A, np.ones(50), 1e-6, 100, 1)
    assert info == 0 and it == ref["iter"]
    assert np.linalg.norm(y - ref["y"]) <= 1e-10 * np.linalg.norm(ref["y"])
    assert np.linalg.norm(y - g["y_exp"]) <= 1e-6 * np.linalg.norm(g["y_exp"])


def test_lanczos_istop2_runs_to_itmax(gpu_ctx):
    A = matgen.laplacian((20, 20), scale=-1.0)
    x = matgen.uniform_pm1(5, 400)
    info, it, err, res, y = _lanczos_gpu(A, x, 1e-10, 8, istop=2)
    ref = O.lanczos(A, x, 1e-10, 8, 2)
    assert info == 0 and it == 7 and err == 0.0
    assert np.linalg.norm(y - ref["y"]) <= 1e-12 * np.linalg.norm(ref["y"])


def test_lanczos_breakdown_and_zero(gpu_ctx):
    m = 30
    A = matgen.laplacian((m,), scale=-1.0)
    x = np.sin(np.arange(1, m + 1) * 3 * np.pi / (m + 1))
    info, it, err, res, y = _lanczos_gpu(A, x, 1e-10, 20)
    ref = O.lanczos(A, x, 1e-10, 20, 1)
    assert info == 0 and it == ref["iter"] == 1
    assert np.linalg.norm(y - ref["y"]) <= 1e-12 * np.linalg.norm(ref["y"])
    info, it, err, res, y = _lanczos_gpu(A, np.zeros(m), 1e-10, 20)
    assert info == 0 and np.all(y == 0.0)


def test_lanczos_errors(gpu_ctx):
    A = matgen.laplacian((6,))
    M = psb.psb_dspmat_type.from_csr(A)
    desc = psb.psb_desc_type(6)
    fun = psfun_d_serial()
    fun.set("VARIANT", "CHBSYM")
    km = psfun_d_krylov()
    km.set("KNAME", "lanczos")       # psb_toupper on kname
    y = _vec(np.zeros(6))
    assert km.apply(fun, M, desc, y, _vec(np.ones(6)), 1e-8, istop=3)[0] == psb.psb_err_invalid_istop_
    assert km.apply(fun, M, desc, psb.psb_d_vect_type(), _vec(np.ones(6)), 1e-8)[0] == \
        psb.psb_err_invalid_vect_state_
    km.set("KNAME", "NOPE")
    assert km.apply(fun, M, desc, y, _vec(np.ones(6)), 1e-8)[0] == psb.psb_err_invalid_args_combination_
    bad = psfun_d_serial()           # default variant 'EXPOKIT' fails inside the f(T) step
    km.set("KNAME", "LANCZOS")
    assert km.apply(bad, M, desc, y, _vec(np.ones(6)), 1e-8)[0] == psb.psb_err_from_subroutine_non_


@pytest.mark.parametrize("fmt", ["AUTO", "HDIA"])
def test_lanczos_full_size_3d_128(gpu_ctx, fmt):
    """128^3 Dirichlet Laplacian (n = 2,097,152), on the automatic choice
    (sliced ELL) and on the diagonal slices the headline runs: against the
    closed-form DST solution and the oracle at 1e-10, iterations within 1."""
    dims = (128, 128, 128)
    A = matgen.laplacian(dims, scale=-1.0)
    x = matgen.uniform_pm1(2024, 128 ** 3)
    info, it, err, res, y = _lanczos_gpu(A, x, 1e-10, 50, fmt=fmt)
    yex = matgen.dst_expm_laplacian(dims, x, 1.0)
    assert info == 0
    assert np.linalg.norm(y - yex) / np.linalg.norm(yex) <= 1e-10
    ref = O.lanczos(A, x, 1e-10, 50, 1)
    assert np.linalg.norm(y - ref["y"]) <= 1e-10 * np.linalg.norm(ref["y"])
    assert abs(it - ref["iter"]) <= 1
    assert len(res) == it + 1 and err < 1e-10


def test_two_contexts_in_two_threads(gpu_ctx):
    """One context per host thread, each on its own stream, running at the
    same time (the ctypes calls release the GIL): every result equals the
    single-context run bitwise (contexts share no state)."""
    import threading
    dims = (40, 40, 40)
    A = matgen.laplacian(dims, scale=-1.0)
    n = 64000
    xs = [matgen.uniform_pm1(50 + i, n) for i in range(2)]
    refs = [_lanczos_gpu(A, x, 1e-10, 50)[4] for x in xs]
    out, errors = {}, []

    def work(i):
        try:
            ctx = psb.Context(0)
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                M = psb.psb_dspmat_type.from_csr(A, ctx)
                fun = psfun_d_serial()
                fun.set("VARIANT", "CHBSYM")
                km = psfun_d_krylov()
                km.set("KNAME", "LANCZOS")
                x = _vec(xs[i])
                ys = []
                for _ in range(6):
                    y = _vec(np.zeros(n))
                    info, it, err, res = km.apply(fun, M, psb.psb_desc_type(n, ctx), y, x, 1e-10, itmax=50)
                    assert info == 0
                    ys.append(y)
                s.synchronize()
                out[i] = [y.get_vect() for y in ys]
            ctx.close()
        except BaseException as e:  # re-raised in the main thread
            errors.append(e)

    th = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errors, errors
    for i in range(2):
        assert len(out[i]) == 6
        for y in out[i]:
            assert np.array_equal(y, refs[i])


def test_objects_destroyed_after_their_context(gpu_ctx):
    """A matrix (with its sliced-ELL and diagonal-slice copies) and a batch
    destroyed after the context they were created on: destroy uses only the
    device they recorded, leaves no HIP error behind, and later calls on
    another context work."""
    from psfun_amd.batched import psfun_batch
    ctx = psb.Context(0)
    A = matgen.laplacian((30, 30))
    M = psb.psb_dspmat_type.from_csr(A, ctx)
    M.set_format("HDIA")
    B = psfun_batch([matgen.rgg_laplacian(300, 5)], ctxt=ctx)
    ctx.close()
    M.free()
    B.free()
    info, it, err, res, y = _lanczos_gpu(matgen.laplacian((30, 30), scale=-1.0), np.ones(900), 1e-10, 30)
    assert info == 0 and it > 1


def test_diagonal_slices_empty_matrix(gpu_ctx):
    """No entries: every slice has no diagonal; the SpMV gives beta*y and the
    Lanczos apply of the zero operator gives y = exp(0) x = x."""
    n = 100
    A = (np.zeros(n + 1, np.int32), np.zeros(0, np.int32), np.zeros(0))
    M = psb.psb_dspmat_type.from_csr(A)
    M.set_format("HDIA")
    assert M.get_format()[0] == "HDIA" and M.hdia_diagonals() == 0
    x = matgen.uniform_pm1(61, n)
    y = _vec(np.full(n, 2.0))
    psb.psb_spmm(1.0, M, _vec(x), 0.5, y, psb.psb_desc_type(n))
    assert np.array_equal(y.get_vect(), np.full(n, 1.0))
    info, it, err, res, yl = _lanczos_gpu(A, x, 1e-10, 30, fmt="HDIA")
    assert info == 0
    assert np.allclose(yl, x, rtol=0, atol=1e-14 * np.linalg.norm(x))


def test_lanczos_itrace_log(gpu_ctx, capsys):
    """itrace = 1 logs every estimate of the run (the reference's drivers
    force it, lanczostest.F90:213-216): one line per iteration with rs(i) for
    itx = i - 1, the last equal to err."""
    from psfun_amd.krylov import lanczos_apply
    A = matgen.laplacian((20, 20), scale=-1.0)
    x = matgen.uniform_pm1(4, 400)
    M = psb.psb_dspmat_type.from_csr(A)
    fun = psfun_d_serial()
    fun.set("VARIANT", "CHBSYM")
    y = psb.psb_d_vect_type(torch.zeros(400, dtype=torch.float64, device="cuda:0"))
    xv = psb.psb_d_vect_type(torch.as_tensor(x, device="cuda:0"))
    capsys.readouterr()
    info, it, err, res = lanczos_apply(fun, M, psb.psb_desc_type(400), y, xv, 1e-10, itmax=60, itrace=1, istop=1)
    lines = capsys.readouterr().out.splitlines()
    assert info == 0 and len(lines) == it + 2
    body = [ln.split(":", 1)[1].split() for ln in lines[1:-1]]
    assert [int(b[0]) for b in body] == list(range(1, it + 1))
    assert np.allclose([float(b[1]) for b in body], res[1:], rtol=1e-8, atol=0)
    assert float(body[-1][1]) == pytest.approx(err, rel=1e-8) and err < 1e-10


@pytest.mark.parametrize("fmt", ["CSRS", "CSR"])
def test_lanczos_power_law_rows(gpu_ctx, fmt):
    """A power-law graph Laplacian (Chung-Lu, rows of 1 to ~3300 entries):
    the CSR row blocks take the rows past a block's 2048 entries alone
    (whole-block sums) and up to 256 short rows per block, against the
    oracle."""
    A = matgen.chung_lu_laplacian(20000, 11)
    d = np.diff(A[0])
    assert d.max() > 2048 and d.min() <= 2
    A = (A[0], A[1], -0.02 * A[2])
    x = matgen.uniform_pm1(12, 20000)
    info, it, err, res, y = _lanczos_gpu(A, x, 1e-10, 120, fmt=fmt)
    ref = O.lanczos(A, x, 1e-10, 120, 1)
    assert info == 0 and abs(it - ref["iter"]) <= 1
    assert np.linalg.norm(y - ref["y"]) <= 1e-10 * np.linalg.norm(ref["y"])


def _arrow(n, seed):
    """Symmetric arrow matrix: diagonal 2 + U[0, 1), first row and column
    U[-1, 1) * 0.5 / sqrt(n): row 0 holds all n entries."""
    d = 2.0 + matgen.uniform01(seed, n)
    a = 0.5 / np.sqrt(n) * matgen.uniform_pm1(seed + 1, n - 1)
    rows = np.concatenate([np.arange(n), np.zeros(n - 1, np.int64), np.arange(1, n)])
    cols = np.concatenate([np.arange(n), np.arange(1, n), np.zeros(n - 1, np.int64)])
    vals = np.concatenate([d, a, a])
    return matgen._coo_to_csr(n, rows, cols, vals)


@pytest.mark.parametrize("n", [9000, 40000])
def test_lanczos_csrs_split_row(gpu_ctx, n):
    """A row longer than the split length (8192 entries) is cut into 4096-entry
    segments summed by different blocks and added in segment order by the
    last to arrive; results against the oracle, and bitwise repeatable."""
    A = _arrow(n, 7)
    A = (A[0], A[1], -A[2])
    x = matgen.uniform_pm1(9, n)
    info, it, err, res, y = _lanczos_gpu(A, x, 1e-10, 60, fmt="CSRS")
    ref = O.lanczos(A, x, 1e-10, 60, 1)
    assert info == 0 and abs(it - ref["iter"]) <= 1
    assert np.linalg.norm(y - ref["y"]) <= 1e-10 * np.linalg.norm(ref["y"])
    info2, it2, err2, res2, y2 = _lanczos_gpu(A, x, 1e-10, 60, fmt="CSRS")
    assert it2 == it and np.array_equal(y2, y) and np.array_equal(res2, res)
