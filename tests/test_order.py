"""Reverse Cuthill-McKee row order of the batched graphs (psfun_rcm_order,
host code of the library; no device needed): a permutation of every graph,
including disconnected and one-directional patterns, whose bandwidth on a
randomly numbered grid and on random geometric graphs comes within 1.3x of
scipy's reverse_cuthill_mckee and far below the random numbering's."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import reverse_cuthill_mckee

from psfun_amd import matgen
from psfun_amd.batched import rcm_order


def _bandwidth(A, perm):
    P = A[perm][:, perm].tocoo()
    return int(np.abs(P.row - P.col).max()) if P.nnz else 0


def _shuffled(A, seed):
    n = A.shape[0]
    p = np.random.default_rng(seed).permutation(n)
    return sp.csr_matrix(A[p][:, p])


@pytest.mark.parametrize("seed", [1, 2])
def test_rcm_grid_bandwidth(seed):
    A = _shuffled(matgen.to_scipy(matgen.laplacian((30, 20), scale=1.0)), seed)
    perm = rcm_order(A.indptr, A.indices)
    assert np.array_equal(np.sort(perm), np.arange(A.shape[0]))
    ours = _bandwidth(A, perm)
    ref = _bandwidth(A, reverse_cuthill_mckee(A, symmetric_mode=True))
    assert ours <= 1.3 * ref
    assert ours <= 0.1 * _bandwidth(A, np.arange(A.shape[0]))


def test_rcm_random_geometric_graph():
    A = matgen.to_scipy(matgen.rgg_laplacian(5000, 11))
    perm = rcm_order(A.indptr, A.indices)
    assert np.array_equal(np.sort(perm), np.arange(5000))
    ours = _bandwidth(A, perm)
    assert ours <= 1.3 * _bandwidth(A, reverse_cuthill_mckee(A, symmetric_mode=True))
    assert ours <= 0.1 * _bandwidth(A, np.arange(5000))


def test_rcm_disconnected_directed_and_offset():
    # two paths, an isolated vertex, and one-directional edges, columns offset by base
    rows = [0, 1, 1, 2, 4, 5, 6, 6]
    cols = [1, 0, 2, 1, 5, 6, 4, 5]
    A = sp.csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(8, 8))
    perm = rcm_order(A.indptr, A.indices + 100, base=100)
    assert np.array_equal(np.sort(perm), np.arange(8))
    assert len(rcm_order(np.zeros(1, np.int32), np.zeros(0, np.int32))) == 0
