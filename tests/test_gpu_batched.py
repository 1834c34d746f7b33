"""GPU parity of the batched Lanczos (many uneven graphs x K vectors) against
the CPU oracle run instance by instance, and against dense expm."""
import numpy as np
import pytest
import scipy.linalg as sla
import torch

from oracle import oracle as O
from psfun_amd import matgen
from psfun_amd.batched import psfun_batch

pytestmark = pytest.mark.gpu


def _graphs(sizes, seed):
    return [matgen.rgg_laplacian(int(n), seed + i) for i, n in enumerate(sizes)]


@pytest.mark.parametrize("K", [1, 3, 6, 16, 64])
def test_batched_matches_oracle_per_instance(gpu_ctx, K):
    sizes = [50, 300, 1000, 37, 2000, 5]
    mats = _graphs(sizes, 100)
    B = psfun_batch(mats)
    N = sum(sizes)
    X = matgen.normal(7, N * K).reshape(N, K)
    t = 0.5
    Xd = torch.as_tensor(X, device="cuda:0")
    Y, iters, errs = B.expm(Xd, t=t, eps=1e-10, itmax=60)
    Y = Y.cpu().numpy()
    off = np.concatenate([[0], np.cumsum(sizes)])
    for g, A in enumerate(mats):
        for k in range(K):
            x = X[off[g]:off[g + 1], k]
            ref = O.lanczos(A, x, 1e-10, 60, 1, fscale=1.0, opscale=-t)
            y = Y[off[g]:off[g + 1], k]
            assert np.linalg.norm(y - ref["y"]) <= 1e-10 * np.linalg.norm(ref["y"]), (g, k)
            assert abs(int(iters[g, k]) - ref["iter"]) <= 1
    # against dense expm where the Krylov space is not exhausted (for the
    # 5-node graph the reference's breakdown rule assembles from T(1:itx-1)
    # and drops the last valid row, so only oracle parity holds there)
    for g in (0, 1):
        Ad = matgen.to_scipy(mats[g]).toarray()
        E = sla.expm(-t * Ad)
        ex = E @ X[off[g]:off[g + 1]]
        assert np.linalg.norm(Y[off[g]:off[g + 1]] - ex) <= 1e-9 * np.linalg.norm(ex)


def test_batched_graph_replay_and_tuning_are_bitwise_stable(gpu_ctx):
    sizes = [400, 1200, 90, 3000]
    B = psfun_batch(_graphs(sizes, 7))
    N = sum(sizes)
    Xd = torch.as_tensor(matgen.normal(3, N * 4).reshape(N, 4), device="cuda:0")
    B.set_tuning(256, False)
    Y0, it0, _ = B.expm(Xd, t=0.5)
    for rpc, g in ((256, True), (1024, True), (64, False)):
        B.set_tuning(rpc, g)
        for _rep in range(3):  # several replays with eager work in between
            Y1, it1, _ = B.expm(Xd, t=0.5)
            torch.cuda.synchronize()
            _ = (Xd * 2.0).sum().item()
            assert np.array_equal(it0, it1)
            if rpc == 256:  # same launch shape: the graph replay is bitwise identical
                assert torch.equal(Y0, Y1)
            else:  # other chunking changes the fold order of the partials
                assert (torch.linalg.norm(Y0 - Y1) / torch.linalg.norm(Y0)).item() <= 1e-10


@pytest.mark.parametrize("K", [2, 4])
def test_batched_zero_column_and_autotune(gpu_ctx, K):
    sizes = [500, 800]
    B = psfun_batch(_graphs(sizes, 11))
    N = sum(sizes)
    X = matgen.normal(5, N * K).reshape(N, K)
    X[:500, 1] = 0.0
    Xd = torch.as_tensor(X, device="cuda:0")
    Y, iters, errs = B.expm(Xd, t=0.5)
    assert torch.all(Y[:500, 1] == 0.0)
    best, table = B.autotune(Xd, t=0.5, candidates=(128, 512), reps=1)
    assert len(table) == 4 and best in [c for c, _ in table]
    Y2, _, _ = B.expm(Xd, t=0.5)
    assert (torch.linalg.norm(Y - Y2) / torch.linalg.norm(Y)).item() <= 1e-10


def _star_path(n, hub, seed):
    """A path of n nodes plus edges from node `hub` to every other node: one
    row of n entries, the rest of 3 to 4 (windows with more entries than an
    LDS stage holds, beside short ones)."""
    rng = np.random.default_rng(seed)
    i = list(range(n - 1)) + [hub] * (n - 1)
    j = list(range(1, n)) + [q for q in range(n) if q != hub]
    keep = [(a, b) for a, b in zip(i, j) if a != b]
    a = np.array([p[0] for p in keep]); b = np.array([p[1] for p in keep])
    key = np.unique(np.minimum(a, b) * n + np.maximum(a, b))
    a, b = key // n, key % n
    w = 0.5 + rng.random(len(a))
    deg = np.bincount(a, weights=w, minlength=n) + np.bincount(b, weights=w, minlength=n)
    rows = np.concatenate([np.arange(n), a, b]); cols = np.concatenate([np.arange(n), b, a])
    vals = np.concatenate([deg, -w, -w])
    import scipy.sparse as sp
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    A.sort_indices()
    return (A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data)


@pytest.mark.parametrize("K", [4, 8])
def test_batched_long_rows_match_oracle(gpu_ctx, K):
    """Graphs whose hub row alone has more entries than a window's LDS stage
    holds, beside graphs of short rows, on the wide-load SpMM (even K) with
    whole and partial 64-row windows: every instance against the oracle."""
    mats = [_star_path(3000, 5, 1), _graphs([700], 3)[0], _star_path(1500, 1400, 2), _graphs([90], 4)[0]]
    sizes = [len(m[0]) - 1 for m in mats]
    B = psfun_batch(mats)
    N = sum(sizes)
    X = matgen.normal(9, N * K).reshape(N, K)
    t = 0.05
    Xd = torch.as_tensor(X, device="cuda:0")
    off = np.concatenate([[0], np.cumsum(sizes)])
    for rpc in (128, 64):
        B.set_tuning(rpc, False)
        Y, iters, errs = B.expm(Xd, t=t, eps=1e-10, itmax=60)
        Y = Y.cpu().numpy()
        for g, A in enumerate(mats):
            for k in range(K):
                ref = O.lanczos(A, X[off[g]:off[g + 1], k], 1e-10, 60, 1, fscale=1.0, opscale=-t)
                y = Y[off[g]:off[g + 1], k]
                assert np.linalg.norm(y - ref["y"]) <= 1e-10 * np.linalg.norm(ref["y"]), (rpc, g, k)
                assert abs(int(iters[g, k]) - ref["iter"]) <= 1


def test_batched_rcm_order_matches_natural(gpu_ctx):
    """A batch kept in reverse Cuthill-McKee order takes and returns X, Y and
    values in the caller's numbering: exp(-tA)X equal to the natural-order
    batch up to the dot products' summation order; SpMM bitwise equal (each
    row keeps its entries in order); set_values takes the caller's entry
    order; calls that would expose the internal numbering refuse."""
    import ctypes as C
    from psfun_amd import _lib
    sizes = [700, 2500, 40, 1800]
    mats = _graphs(sizes, 31)
    Bn = psfun_batch(mats)
    Br = psfun_batch(mats, order="rcm")
    N, K = sum(sizes), 5
    Xd = torch.as_tensor(matgen.normal(5, N * K).reshape(N, K), device="cuda:0")
    Yn, itn, _ = Bn.expm(Xd, t=0.5)
    Yr, itr, _ = Br.expm(Xd, t=0.5)
    assert np.abs(itn - itr).max() <= 1
    assert torch.linalg.norm(Yr - Yn) <= 1e-12 * torch.linalg.norm(Yn)
    for rpc, g in ((256, True), (1024, False)):        # tuning and graph replay keep the permutation
        Br.set_tuning(rpc, g)                           # (other chunking: other fold order of the partials)
        Y2, _, _ = Br.expm(Xd, t=0.5)
        assert torch.linalg.norm(Y2 - Yr) <= 1e-10 * torch.linalg.norm(Yr)
    L = _lib.lib()
    Y0 = torch.as_tensor(matgen.normal(6, N * K).reshape(N, K), device="cuda:0")
    Za, Zb = Y0.clone(), Y0.clone()
    for B, Z in ((Bn, Za), (Br, Zb)):
        B.ctxt.bind_stream()
        _lib.check(L.psfun_batch_spmm(B.ctxt.handle, B.handle, K, -0.5, C.c_void_p(Xd.data_ptr()), 0.25,
                                      C.c_void_p(Z.data_ptr())))
    torch.cuda.synchronize()
    assert torch.equal(Za, Zb)
    vals = torch.as_tensor(np.concatenate([m[2] for m in mats]) * 1.7, device="cuda:0")
    for B in (Bn, Br):
        _lib.check(L.psfun_batch_set_values(B.handle, C.c_void_p(vals.data_ptr())))
    Yn, _, _ = Bn.expm(Xd, t=0.5)
    Yr, _, _ = Br.expm(Xd, t=0.5)
    assert torch.linalg.norm(Yr - Yn) <= 1e-12 * torch.linalg.norm(Yn)
    rp = C.c_void_p()
    assert L.psfun_batch_csr(Br.handle, C.byref(rp), None, None) == _lib.PSFUN_ERR_UNSUPPORTED


def test_exported_basis_is_unnormalised_and_orthonormal_after_scaling(gpu_ctx):
    """psfun_batch_lanczos_basis keeps the basis unnormalised: slot 0 holds x
    bitwise, and v_{l+1} = slot l / T_beta(l) (v_1 = x / rs1) are orthonormal
    to rounding over a short run; y equals the plain batched result."""
    from psfun_amd.autograd import ExpmvOperator
    A = matgen.laplacian((30, 30))
    n, K = 900, 3
    op = ExpmvOperator(A[0], A[1], eps=1e-10, itmax=40)
    op.set_values(torch.as_tensor(A[2], device="cuda:0"))
    X = torch.as_tensor(matgen.normal(41, n * K).reshape(n, K), device="cuda:0")
    Y, (V, Ta, Tb, nb, rs1, K_) = op.expm_basis(X, 0.05)
    Yp = op.expm(X, 0.05)
    assert torch.equal(Y, Yp)
    ld = op.itmax + 2
    V = V.cpu().numpy().reshape(op.itmax - 1, n, K)
    Tb = Tb.cpu().numpy().reshape(K, ld)
    nb = nb.cpu().numpy()
    rs1 = rs1.cpu().numpy()
    Xh = X.cpu().numpy()
    for k in range(K):
        m = int(nb[k])
        assert m >= 5
        assert np.array_equal(V[0, :, k], Xh[:, k])
        Q = np.stack([V[l, :, k] / (rs1[k] if l == 0 else Tb[k, l]) for l in range(m)], axis=1)
        assert np.max(np.abs(Q.T @ Q - np.eye(m))) <= 1e-10
