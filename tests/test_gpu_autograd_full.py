"""Differentiable exp(-tA)X at the full anisotropic workload: the 1024 x 1024
2-D operator (log-normal conductances, sigma 1, x-edges x10; n = 1 048 576,
nnz = 5 238 784), t = 0.1, normal columns X and G (4 columns here; the bench
runs 32 with the same per-column algorithm).

- Y and dL/dX against the oracle's restatement of
  krylov/polynomial/psfun_d_lanczos.f90:183-377 per column: relative 2-norm
  <= 1e-10.
- dL/dt against -<G, A Y> formed on the host (scipy), relative 1e-10.
- dL/d(nonzeros) against the multi-shift CG mode (the same Talbot rule with
  full shifted solves instead of Galerkin resolvents from the bases),
  relative 2-norm <= 1e-8, and against a central difference along a random
  symmetric direction (h = 1e-6, which resolves the derivative to about
  eps_forward*|L|/h, hence 5e-5).
- The one-call C-ABI VJP (psfun_expv_vjp) bitwise equal to autograd."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle import oracle as O
from psfun_amd import matgen
from psfun_amd.autograd import ExpmvOperator, expmv

pytestmark = pytest.mark.gpu

M, T, K, EPS, ITMAX, POLES = 1024, 0.1, 4, 1e-10, 200, 16


@pytest.fixture(scope="module")
def aniso():
    print(f"\n[adjoint-full] operator {M}^2, K={K}", flush=True)
    A = matgen.anisotropic_2d(M, seed=11)
    n = M * M
    X = matgen.normal(21, n * K).reshape(n, K)
    G = matgen.normal(22, n * K).reshape(n, K)
    off = np.array([0, n], np.int64)
    print("[adjoint-full] oracle forward and adjoint", flush=True)
    Yref, _, _ = O.lanczos_batch(off, A, X, EPS, ITMAX, 1, opscale=-T)
    gXref, _, _ = O.lanczos_batch(off, A, G, EPS, ITMAX, 1, opscale=-T)
    return A, X, G, Yref, gXref


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_full_size_gradients(gpu_ctx, aniso):
    A, X, G, Yref, gXref = aniso
    rp, ci, va = A
    dev = "cuda:0"
    op = ExpmvOperator(rp, ci, eps=EPS, itmax=ITMAX, quad_poles=POLES, store_basis=True)
    vals = torch.tensor(va, device=dev, requires_grad=True)
    Xd = torch.tensor(X, device=dev, requires_grad=True)
    t = torch.tensor(T, dtype=torch.float64, device=dev, requires_grad=True)
    Gd = torch.tensor(G, device=dev)
    print("[adjoint-full] krylov forward + backward", flush=True)
    Y = expmv(vals, Xd, t, op)
    (Gd * Y).sum().backward()
    Yh = Y.detach().cpu().numpy()
    assert _rel(Yh, Yref) <= 1e-10
    assert _rel(Xd.grad.cpu().numpy(), gXref) <= 1e-10
    S = matgen.to_scipy(A)
    gt = -float(np.sum(G * (S @ Yh)))
    assert abs(t.grad.item() - gt) <= 1e-10 * abs(gt)
    gk = vals.grad.cpu().numpy()

    print("[adjoint-full] one-call VJP", flush=True)
    with torch.no_grad():
        op.set_values(vals)
        gX1, gt1, gv1 = op.vjp(Xd.detach(), Y.detach(), Gd, T)
    assert torch.equal(gX1, Xd.grad)
    assert torch.equal(gv1, vals.grad)
    assert abs(gt1 - t.grad.item()) <= 1e-12 * abs(gt)

    print("[adjoint-full] multi-shift CG mode", flush=True)
    op2 = ExpmvOperator(rp, ci, eps=EPS, itmax=ITMAX, quad_poles=POLES, quad_eps=1e-13, grad_mode="mscg")
    vals2 = torch.tensor(va, device=dev, requires_grad=True)
    Y2 = expmv(vals2, torch.tensor(X, device=dev), torch.tensor(T, dtype=torch.float64, device=dev), op2)
    (Gd * Y2).sum().backward()
    assert _rel(gk, vals2.grad.cpu().numpy()) <= 1e-8

    print("[adjoint-full] central difference", flush=True)
    n = M * M
    rng = np.random.default_rng(5)
    D = sp.csr_matrix((rng.normal(size=len(va)), ci, rp), shape=(n, n))
    D = ((D + D.T) * 0.5).tocsr()
    D.sort_indices()
    assert np.array_equal(D.indptr, rp) and np.array_equal(D.indices, ci)
    d = torch.tensor(D.data, device=dev)
    h = 1e-6
    with torch.no_grad():
        Xc, tc = Xd.detach(), t.detach()
        lp = float((Gd * expmv(vals.detach() + h * d, Xc, tc, op)).sum().item())
        lm = float((Gd * expmv(vals.detach() - h * d, Xc, tc, op)).sum().item())
    fd = (lp - lm) / (2 * h)
    an = float(np.dot(gk, D.data))
    assert abs(fd - an) <= 5e-5 * max(abs(an), 1.0)
