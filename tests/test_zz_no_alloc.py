"""psfun_lanczos_apply_noalloc: the Lanczos apply on the library's fixed
device scratch, called through the C-ABI.  Its results equal
psfun_lanczos_apply's bitwise (same kernels, same launch shapes, same fold
orders) on two calls in a row, the second with a smaller input that reuses
the scratch the first left behind; it allocates no device memory; inputs
past the 64 MiB scratch are refused before any work."""
import ctypes as C

import numpy as np
import pytest
import torch

from psfun_amd import _lib, matgen, psb

pytestmark = pytest.mark.gpu


def _apply(L, fn, ctx, M, x, y, itmax=50, eps=1e-10, istop=1):
    it, nres = C.c_int(0), C.c_int(0)
    err = C.c_double(0.0)
    res = np.zeros(itmax + 2)
    st = fn(ctx.handle, M.handle, _lib.PSFUN_FN_EXP, 1.0, 1.0, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()),
            eps, itmax, istop, C.byref(it), C.byref(err), res.ctypes.data_as(C.POINTER(C.c_double)), C.byref(nres))
    torch.cuda.synchronize()
    return st, it.value, err.value, res[:nres.value].copy()


def test_noalloc_matches_allocating_apply_twice(gpu_ctx):
    L = _lib.lib()
    ctx = psb.Context(0)            # fresh context: its own workspace is never grown below
    ctx.bind_stream()
    cases = [((24, 24, 24), 50, "HDIA"), ((100, 100), 30, "SELL")]   # the second call is smaller (1e4 < 13824 rows)
    mats = []
    for dims, itmax, fmt in cases:
        A = matgen.laplacian(dims, scale=-1.0)
        M = psb.psb_dspmat_type.from_csr(A, ctx)
        M.set_format(fmt)
        n = int(np.prod(dims))
        x = torch.as_tensor(matgen.uniform_pm1(5 + n, n), device="cuda:0")
        mats.append((dims, itmax, M, x, n))
    torch.cuda.synchronize()
    ys = [torch.full((n,), np.nan, dtype=torch.float64, device="cuda:0") for *_, n in mats]
    free0 = torch.cuda.mem_get_info(0)[0]
    out = []
    for (dims, itmax, M, x, n), y in zip(mats, ys):
        print(f"[noalloc] dims={dims} itmax={itmax} format={M.get_format()[0]}", flush=True)
        out.append(_apply(L, L.psfun_lanczos_apply_noalloc, ctx, M, x, y, itmax=itmax))
    free1 = torch.cuda.mem_get_info(0)[0]
    # no device memory was taken by the two calls: the 13824 x 50 basis alone
    # would be 5.5 MB of workspace (the runtime's own bookkeeping may move a page)
    assert free0 - free1 < (2 << 20), (free0, free1)
    for (dims, itmax, M, x, n), y, (st, it, err, res) in zip(mats, ys, out):
        assert st == _lib.PSFUN_SUCCESS, _lib.lib().psfun_strerror(st)
        yref = torch.zeros(n, dtype=torch.float64, device="cuda:0")
        st2, it2, err2, res2 = _apply(L, L.psfun_lanczos_apply, ctx, M, x, yref, itmax=itmax)
        assert st2 == _lib.PSFUN_SUCCESS
        assert torch.equal(y, yref) and it == it2 and err == err2 and np.array_equal(res, res2)
        yex = matgen.dst_expm_laplacian(dims, x.cpu().numpy(), 1.0)
        assert np.linalg.norm(y.cpu().numpy() - yex) <= 1e-10 * np.linalg.norm(yex)
    for *_, M, x, n in mats:
        M.free()
    ctx.close()


def test_noalloc_refuses_past_scratch_and_userf(gpu_ctx):
    L = _lib.lib()
    ctx = psb.default_context(0)
    ctx.bind_stream()
    n = 200_000                       # 8 * 200 000 * 50 bytes of basis > 64 MiB
    A = matgen.laplacian((n,), scale=-1.0)
    M = psb.psb_dspmat_type.from_csr(A, ctx)
    x = torch.ones(n, dtype=torch.float64, device="cuda:0")
    y = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    st, *_ = _apply(L, L.psfun_lanczos_apply_noalloc, ctx, M, x, y, itmax=50)
    assert st == _lib.PSFUN_ERR_TOO_LARGE
    assert b"64 MiB" in L.psfun_strerror(st)
    assert torch.all(y == 0.0)        # refused before any work
    st, it, err, res = _apply(L, L.psfun_lanczos_apply_noalloc, ctx, M, x, y, itmax=20)   # fits at itmax 20
    assert st == _lib.PSFUN_SUCCESS and it >= 1
    it_, nres = C.c_int(0), C.c_int(0)
    err_ = C.c_double(0.0)
    st = L.psfun_lanczos_apply_noalloc(ctx.handle, M.handle, _lib.PSFUN_FN_USERF, 1.0, 1.0, C.c_void_p(x.data_ptr()),
                                       C.c_void_p(y.data_ptr()), 1e-10, 20, 1, C.byref(it_), C.byref(err_), None,
                                       C.byref(nres))
    assert st == _lib.PSFUN_ERR_UNSUPPORTED
