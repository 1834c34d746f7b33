"""A sweep of operators, time scales and signs through the device Lanczos,
batched Lanczos and Arnoldi, each against the oracle's restatement of
psfun_d_lanczos.f90:183-377 / psfun_d_arnoldi.f90:183-329 and against
scipy's expm_multiply:

- operators: 2-D and 3-D Dirichlet Laplacians, the anisotropic diffusion
  operator (log-normal conductances, x-edges x10), a random geometric graph
  Laplacian (singular), a random sparse SPD matrix and a random sparse
  symmetric indefinite matrix;
- exp(-tA)x for t from 0.01 to 3 (||tA|| from ~0.1 to ~800, up to 106
  Lanczos steps) and exp(+tA)x for small t (growth).

Bars: relative 2-norm 1e-9 against the oracle and scipy, iterations within 2
(long runs lose orthogonality, and the last digits of the stopping estimate
then follow each summation order)."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch
from scipy.sparse.linalg import expm_multiply

from oracle import oracle as O
from psfun_amd import matgen, psb
from psfun_amd.arnoldi import arnoldi_apply
from psfun_amd.batched import psfun_batch
from psfun_amd.krylov import lanczos_apply
from psfun_amd.serial import psfun_d_serial

pytestmark = pytest.mark.gpu

EPS, ITMAX = 1e-10, 250


def _csr(S):
    S = sp.csr_matrix(S)
    S.sum_duplicates()
    S.sort_indices()
    return S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data.astype(np.float64)


def _random_spd(n, seed):
    rng = np.random.default_rng(seed)
    R = sp.random(n, n, density=6.0 / n, random_state=rng, format="csr")
    S = R + R.T
    d = np.asarray(abs(S).sum(1)).ravel() + rng.uniform(0.1, 5.0, n)
    return _csr(S + sp.diags(d))


def _random_indefinite(n, seed):
    rng = np.random.default_rng(seed)
    R = sp.random(n, n, density=6.0 / n, random_state=rng, format="csr")
    return _csr(R + R.T + sp.diags(rng.uniform(-3.0, 3.0, n)))


OPERATORS = {
    "lap2d": lambda: matgen.laplacian((60, 50)),
    "lap3d": lambda: matgen.laplacian((16, 14, 12)),
    "aniso48": lambda: matgen.anisotropic_2d(48, seed=3),
    "rgg": lambda: matgen.rgg_laplacian(3000, 7),
    "spd": lambda: _random_spd(2500, 5),
    "indef": lambda: _random_indefinite(2500, 6),
}
CASES = [("lap2d", -0.5), ("lap2d", -3.0), ("lap2d", 0.05), ("lap3d", -1.0), ("aniso48", -0.01),
         ("aniso48", -0.3), ("aniso48", -1.0), ("rgg", -0.5), ("rgg", -2.0), ("spd", -0.2), ("spd", -1.0),
         ("indef", -0.2), ("indef", 0.2)]
_cache = {}


def _case(name, opscale):
    key = (name, opscale)
    if key not in _cache:
        A = OPERATORS[name]()
        n = len(A[0]) - 1
        x = matgen.normal(3 + sorted(OPERATORS).index(name), n)
        ys = expm_multiply(opscale * matgen.to_scipy(A), x)
        _cache[key] = (A, x, ys)
    return _cache[key]


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _vec(v):
    return psb.psb_d_vect_type(torch.as_tensor(np.ascontiguousarray(v), device="cuda:0"))


@pytest.mark.parametrize("name,opscale", CASES)
def test_lanczos_sweep(gpu_ctx, name, opscale):
    A, x, ys = _case(name, opscale)
    n = len(x)
    ref = O.lanczos(A, x, EPS, ITMAX, 1, opscale=opscale)
    assert _rel(ref["y"], ys) <= 1e-9                 # the oracle itself
    M = psb.psb_dspmat_type.from_csr(A)
    fun = psfun_d_serial()
    fun.set("VARIANT", "CHBSYM")
    y = _vec(np.zeros(n))
    info, it, err, res = lanczos_apply(fun, M, psb.psb_desc_type(n), y, _vec(x), EPS, itmax=ITMAX,
                                       opscale=opscale)
    assert info == 0
    yg = y.get_vect()
    print(f"[stress] lanczos {name} opscale={opscale} it={it}/{ref['iter']} "
          f"oracle {_rel(yg, ref['y']):.2e} scipy {_rel(yg, ys):.2e}", flush=True)
    assert abs(it - ref["iter"]) <= 2
    assert _rel(yg, ref["y"]) <= 1e-9
    assert _rel(yg, ys) <= 1e-9


@pytest.mark.parametrize("name,opscale", CASES)
def test_batched_sweep(gpu_ctx, name, opscale):
    A, x, ys = _case(name, opscale)
    n = len(x)
    X = np.stack([x, matgen.uniform_pm1(9, n)], 1)
    B = psfun_batch([A])
    Xd = torch.as_tensor(X, device="cuda:0")
    Yd = torch.empty_like(Xd)
    iters, _ = B.lanczos(Xd, Yd, fscale=1.0, opscale=opscale, eps=EPS, itmax=ITMAX)
    Yh = Yd.cpu().numpy()
    for k in range(2):
        ref = O.lanczos(A, X[:, k].copy(), EPS, ITMAX, 1, opscale=opscale)
        assert abs(int(iters[0, k]) - ref["iter"]) <= 2
        assert _rel(Yh[:, k], ref["y"]) <= 1e-9
    assert _rel(Yh[:, 0], ys) <= 1e-9


@pytest.mark.parametrize("name,opscale", CASES)
def test_arnoldi_sweep(gpu_ctx, name, opscale):
    A, x, ys = _case(name, opscale)
    n = len(x)
    ref = O.arnoldi(A, x, EPS, ITMAX, 1, scaling=opscale)
    M = psb.psb_dspmat_type.from_csr(A)
    fun = psfun_d_serial()
    fun.set("VARIANT", "CHBHES")
    fun.set("SCALING", opscale)
    y = _vec(np.zeros(n))
    info, it, err, res = arnoldi_apply(fun, M, psb.psb_desc_type(n), y, _vec(x), EPS, itmax=ITMAX)
    assert info == 0
    yg = y.get_vect()
    assert abs(it - ref["iter"]) <= 2
    assert _rel(yg, ref["y"]) <= 1e-9
    assert _rel(yg, ys) <= 1e-9
