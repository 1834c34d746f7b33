"""GPU checks of the differentiable exp(-tA)X: against the dense FP64
Frechet-derivative oracle on a 32x32-grid anisotropic operator (fixture) and
against central finite differences (step 1e-6).  A central difference with
h = 1e-6 resolves a derivative only to about eps_forward*|L|/h (forward
tolerance 1e-12, |L| ~ 10-100) plus h^2 terms, hence its 5e-5 bound; the
oracle comparison carries the tight bounds."""
import numpy as np
import pytest
import torch

from conftest import load_golden
from psfun_amd.autograd import ExpmvOperator, expmv

pytestmark = pytest.mark.gpu


def _setup(mode="krylov", store=True):
    g = load_golden("frechet_small.npz")
    op = ExpmvOperator(g["rowptr"], g["colind"], eps=1e-12, itmax=200, quad_poles=24, quad_eps=1e-14,
                       store_basis=store, grad_mode=mode)
    dev = "cuda:0"
    vals = torch.tensor(g["val"], device=dev, requires_grad=True)
    X = torch.tensor(g["X"], device=dev, requires_grad=True)
    t = torch.tensor(float(g["t"]), dtype=torch.float64, device=dev, requires_grad=True)
    G = torch.tensor(g["G"], device=dev)
    return g, op, vals, X, t, G


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("mode,store", [("krylov", True), ("krylov", False), ("krylov", "auto"), ("mscg", True)])
def test_forward_and_gradients_vs_frechet_oracle(gpu_ctx, mode, store):
    g, op, vals, X, t, G = _setup(mode, store)
    Y = expmv(vals, X, t, op)
    assert _rel(Y.detach().cpu().numpy(), g["Y"]) <= 1e-10
    loss = (G * Y).sum()
    loss.backward()
    assert _rel(X.grad.cpu().numpy(), g["grad_X"]) <= 1e-10
    assert abs(t.grad.item() - float(g["grad_t"])) <= 1e-10 * abs(float(g["grad_t"]))
    assert _rel(vals.grad.cpu().numpy(), g["grad_val"]) <= 1e-9


def test_gradients_vs_central_differences(gpu_ctx):
    g, op, vals, X, t, G = _setup()
    Y = expmv(vals, X, t, op)
    (G * Y).sum().backward()

    def loss(v, x, tt):
        with torch.no_grad():
            return float((G * expmv(v, x, tt, op)).sum().item())

    h = 1e-6
    rng = np.random.default_rng(5)
    with torch.no_grad():
        # symmetric direction in the nonzeros of A (the operator is symmetric;
        # a one-sided perturbation would leave the Lanczos forward's domain)
        import scipy.sparse as sp
        n = len(g["rowptr"]) - 1
        D = sp.csr_matrix((rng.normal(size=vals.numel()), g["colind"], g["rowptr"]), shape=(n, n))
        D = ((D + D.T) * 0.5).tocsr()
        D.sort_indices()
        d = torch.tensor(D.data, device=vals.device)
        fd = (loss(vals + h * d, X, t) - loss(vals - h * d, X, t)) / (2 * h)
        an = float((vals.grad * d).sum().item())
        assert abs(fd - an) <= 5e-5 * max(abs(an), 1.0)
        # direction in X
        dX = torch.tensor(rng.normal(size=tuple(X.shape)), device=X.device)
        fd = (loss(vals, X + h * dX, t) - loss(vals, X - h * dX, t)) / (2 * h)
        an = float((X.grad * dX).sum().item())
        assert abs(fd - an) <= 5e-5 * max(abs(an), 1.0)
        # t
        fd = (loss(vals, X, t + h) - loss(vals, X, t - h)) / (2 * h)
        assert abs(fd - t.grad.item()) <= 5e-5 * max(abs(t.grad.item()), 1.0)


def test_basis_memory_decision(gpu_ctx):
    """store_basis="auto": keep the basis when two bases fit in half the free
    device memory, recompute otherwise (checked with a pattern whose basis
    cannot fit and one whose basis does)."""
    g = load_golden("frechet_small.npz")
    op = ExpmvOperator(g["rowptr"], g["colind"], itmax=200)
    assert op.keep_basis(2, "cuda:0")
    free, _ = torch.cuda.mem_get_info("cuda:0")
    K_huge = int(free // op.basis_bytes(1)) + 1
    assert not op.keep_basis(K_huge, "cuda:0")



def test_one_call_vjp_matches_autograd_and_oracle(gpu_ctx):
    """psfun_expv_vjp (C-ABI, forward basis recomputed) against the autograd
    backward that recomputes its basis: the same kernels in the same order, so
    dL/dX and dL/dvals are bitwise equal; dL/dt differs only in the order of
    its dot product.  Both against the dense Frechet fixture."""
    g, op, vals, X, t, G = _setup("krylov", False)
    Y = expmv(vals, X, t, op)
    (G * Y).sum().backward()
    with torch.no_grad():
        op.set_values(vals)
        gX, gt, gv = op.vjp(X.detach(), Y.detach(), G, float(t))
    assert torch.equal(gX, X.grad)
    assert torch.equal(gv, vals.grad)
    assert abs(gt - t.grad.item()) <= 1e-13 * abs(t.grad.item())
    assert _rel(gX.cpu().numpy(), g["grad_X"]) <= 1e-10
    assert _rel(gv.cpu().numpy(), g["grad_val"]) <= 1e-9
    assert abs(gt - float(g["grad_t"])) <= 1e-10 * abs(float(g["grad_t"]))
    # partial requests: each output alone equals the full call's
    with torch.no_grad():
        gX2, gt2, gv2 = op.vjp(X.detach(), Y.detach(), G, float(t), need_t=False, need_a=False)
        assert gt2 is None and gv2 is None and torch.equal(gX2, gX)
        gX3, gt3, gv3 = op.vjp(X.detach(), Y.detach(), G, float(t), need_x=False, need_t=False)
        assert gX3 is None and gt3 is None and torch.equal(gv3, gv)


def test_one_call_vjp_argument_errors(gpu_ctx):
    import ctypes as C
    from psfun_amd import _lib
    g, op, vals, X, t, G = _setup("krylov", False)
    L = _lib.lib()
    x = X.detach().contiguous()
    gv = torch.empty(op.nnz, dtype=torch.float64, device=x.device)
    gt = C.c_double(0.0)
    p = lambda a: C.c_void_p(a.data_ptr())
    K = int(x.shape[1])
    h, b = op.ctxt.handle, op.batch.handle
    # dL/dt without Y, dL/dvals without X or with t <= 0, K out of range
    assert L.psfun_expv_vjp(h, b, K, 0.1, p(x), None, p(G), 1e-12, 200, 24, None, C.byref(gt), None) == \
        _lib.PSFUN_ERR_INVALID_ARG
    assert L.psfun_expv_vjp(h, b, K, 0.1, None, p(x), p(G), 1e-12, 200, 24, None, None, p(gv)) == \
        _lib.PSFUN_ERR_INVALID_ARG
    assert L.psfun_expv_vjp(h, b, K, 0.0, p(x), p(x), p(G), 1e-12, 200, 24, None, None, p(gv)) == \
        _lib.PSFUN_ERR_INVALID_ARG
    assert L.psfun_expv_vjp(h, b, 65, 0.1, p(x), p(x), p(G), 1e-12, 200, 24, None, None, p(gv)) == \
        _lib.PSFUN_ERR_INVALID_ARG


@pytest.mark.parametrize("nrows,mmax", [(5000, 13), (40, 4)])
def test_resolvent_kernels_agree_bitwise(gpu_ctx, nrows, mmax):
    """psfun_krylov_resolvents takes k_krylov_resolvents_rows<8, 4> (four
    rows per thread) for up to 8 shifts and <16, 2> past 8: for the same
    data, the first 8 shifts' resolvents of a 9-shift call equal the 8-shift
    call's bitwise (each element's sum runs over l and the shifts in the
    same order).  A partial last tile, a basis length that is not a multiple
    of the coefficient chunk, and a column with a shorter basis."""
    import ctypes as C
    from psfun_amd import _lib
    K = 32
    rng = np.random.default_rng(3)
    dev = "cuda:0"
    V = torch.as_tensor(rng.standard_normal((mmax, nrows, K)), device=dev)
    ldt = mmax + 2
    Ta = torch.as_tensor(np.concatenate([np.zeros((K, 1)), 1.0 + 2.0 * rng.random((K, ldt - 1))], 1), device=dev)
    Tb = torch.as_tensor(np.concatenate([np.zeros((K, 1)), 0.1 + 0.4 * rng.random((K, ldt - 1))], 1), device=dev)
    nb = np.full(K, mmax, np.int32)
    nb[5] = max(1, mmax - 3)
    nbd = torch.as_tensor(nb, device=dev)
    rs1 = torch.as_tensor(1.0 + rng.random(K), device=dev)
    sh = np.stack([0.5 + rng.random(9), rng.standard_normal(9)], 1).reshape(-1)
    outs = []
    ctx = _lib.lib()
    from psfun_amd import psb
    h = psb.default_context(0)
    for ns in (8, 9):
        out = torch.empty((ns, nrows, K), dtype=torch.complex128, device=dev)
        h.bind_stream()
        st = ctx.psfun_krylov_resolvents(h.handle, nrows, K, mmax, C.c_void_p(V.data_ptr()), C.c_void_p(Ta.data_ptr()),
                                         C.c_void_p(Tb.data_ptr()), ldt, 0.5, C.c_void_p(nbd.data_ptr()),
                                         C.c_void_p(rs1.data_ptr()), ns, sh.ctypes.data_as(C.c_void_p),
                                         C.c_void_p(out.data_ptr()))
        _lib.check(st, "psfun_krylov_resolvents")
        torch.cuda.synchronize()
        outs.append(out.cpu().numpy())
    assert np.array_equal(outs[0], outs[1][:8])
    assert np.all(np.isfinite(outs[0])) and np.abs(outs[0]).max() > 0
