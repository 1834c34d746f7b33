"""Batched Lanczos parity at the full batched heat-kernel workload: 256
random-geometric-graph Laplacians (n log-uniform in [1k, 64k], mean degree
about 8, weights U[0.5, 1.5]; 3.7 M rows, 33 M nonzeros) x 16 normal
vectors, exp(-0.5 L) X, eps 1e-10.  Every one of the 4096 (graph, column)
instances is compared with the oracle's restatement of
krylov/polynomial/psfun_d_lanczos.f90:183-377 run on that instance alone:
relative 2-norm <= 1e-10 and iterations within 1, in the graphs' own row
order and in reverse Cuthill-McKee order, eagerly and replayed from a
hipGraph, at the chunk sizes the autotuner chooses between."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from psfun_amd import matgen
from psfun_amd.batched import concat_block_diagonal, psfun_batch

pytestmark = pytest.mark.gpu

T, K, EPS, ITMAX = 0.5, 16, 1e-10, 60


@pytest.fixture(scope="module")
def rgg_batch():
    sizes = matgen.rgg_sizes(256, 3)
    print(f"\n[config3] generating 256 graphs, sum n = {int(sizes.sum())}", flush=True)
    mats = [matgen.rgg_laplacian(int(n), 1000 + i) for i, n in enumerate(sizes)]
    off, rp, ci, va = concat_block_diagonal(mats)
    N = int(off[-1])
    X = matgen.normal(77, N * K).reshape(N, K)
    print(f"[config3] oracle on {len(mats) * K} instances", flush=True)
    Yref, itref, _ = O.lanczos_batch(off, (rp, ci, va), X, EPS, ITMAX, 1, opscale=-T)
    nref = np.sqrt(np.add.reduceat(Yref * Yref, off[:-1], axis=0))
    return mats, sizes, off, X, Yref, itref, nref


def _per_instance_rel(Y, Yref, off, nref):
    d = Y - Yref
    return np.sqrt(np.add.reduceat(d * d, off[:-1], axis=0)) / nref


@pytest.mark.parametrize("order,use_graph,rows_per_chunk",
                         [("natural", False, 512), ("natural", True, 256),
                          ("rcm", False, 1024), ("rcm", True, 256)])
def test_rgg256x16_matches_oracle_per_instance(gpu_ctx, rgg_batch, order, use_graph, rows_per_chunk):
    mats, sizes, off, X, Yref, itref, nref = rgg_batch
    assert sizes.min() < 1100 and sizes.max() > 55000    # the smallest and largest graphs are in the batch
    print(f"[config3] order={order} graph={use_graph} rows_per_chunk={rows_per_chunk}", flush=True)
    B = psfun_batch(mats, order=order)
    B.set_tuning(rows_per_chunk, use_graph)
    Xd = torch.as_tensor(X, device="cuda:0")
    Y, iters, errs = B.expm(Xd, t=T, eps=EPS, itmax=ITMAX)
    Yh = Y.cpu().numpy()
    rel = _per_instance_rel(Yh, Yref, off, nref)
    assert rel.shape == (256, K)
    worst = np.unravel_index(np.argmax(rel), rel.shape)
    assert rel.max() <= 1e-10, (worst, rel.max())
    assert np.abs(iters - itref).max() <= 1
    assert np.all(errs < EPS)
    if use_graph:
        # replays with eager work and synchronisations in between: bitwise the first run
        for _ in range(2):
            _ = (Xd * 2.0).sum().item()
            Y2, it2, _ = B.expm(Xd, t=T, eps=EPS, itmax=ITMAX)
            torch.cuda.synchronize()
            assert torch.equal(Y2, Y) and np.array_equal(it2, iters)
    B.free()
