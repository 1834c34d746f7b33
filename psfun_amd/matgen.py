"""Seeded synthetic inputs for the f(A)x workloads.

* SplitMix64 -> U[-1, 1) and Box-Muller normals (identical definitions are
  used wherever the same inputs are regenerated).
* Dirichlet Laplacians (2-D 5-point, 3-D 7-point, unscaled: diagonal 2d,
  off-diagonals -1), the 2-D anisotropic diffusion operator with log-normal
  edge conductances, and random geometric graph Laplacians L = D - W.
* The closed-form DST solution exp(-tA)x for the Dirichlet Laplacians:
  A = sum_d I x .. x L1 x .. x I with L1 = tridiag(-1, 2, -1) of order m has
  the orthonormal DST-I eigenvectors and eigenvalues 2 - 2 cos(k pi/(m+1)).

Matrices are returned as (rowptr, colind, val) int32/int32/float64 arrays,
0-based, columns sorted within each row.
"""
import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, count, offset=0):
    """`count` consecutive SplitMix64 outputs starting after `offset` draws."""
    idx = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + idx * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, count, offset=0):
    return (splitmix64(seed, count, offset) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def uniform_pm1(seed, count, offset=0):
    """U[-1, 1): 2u - 1."""
    return 2.0 * uniform01(seed, count, offset) - 1.0


def normal(seed, count, offset=0):
    """Box-Muller on consecutive pairs (u1, u2): sqrt(-2 ln(1-u1)) cos(2 pi u2)
    and ... sin(2 pi u2)."""
    npair = (count + 1) // 2
    u = uniform01(seed, 2 * npair, 2 * offset)
    u1, u2 = u[0::2], u[1::2]
    rad = np.sqrt(-2.0 * np.log1p(-u1))
    z = np.empty(2 * npair)
    z[0::2] = rad * np.cos(2 * np.pi * u2)
    z[1::2] = rad * np.sin(2 * np.pi * u2)
    return z[:count]


def _coo_to_csr(n, rows, cols, vals):
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    counts = np.bincount(rows, minlength=n)
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=rowptr[1:])
    return rowptr.astype(np.int32), cols.astype(np.int32), np.ascontiguousarray(vals, dtype=np.float64)


def laplacian(dims, scale=1.0):
    """Dirichlet Laplacian on a grid of shape `dims` (1, 2 or 3 axes),
    natural ordering with the first axis fastest; scale*A is returned."""
    dims = tuple(int(d) for d in dims)
    n = int(np.prod(dims))
    idx = np.arange(n, dtype=np.int64)
    coords = []
    rem = idx.copy()
    for d in dims:
        coords.append(rem % d)
        rem //= d
    rows = [idx]
    cols = [idx]
    vals = [np.full(n, 2.0 * len(dims) * scale)]
    stride = 1
    for ax, d in enumerate(dims):
        c = coords[ax]
        for step in (-1, 1):
            ok = (c + step >= 0) & (c + step < d)
            rows.append(idx[ok])
            cols.append(idx[ok] + step * stride)
            vals.append(np.full(int(ok.sum()), -1.0 * scale))
        stride *= d
    return _coo_to_csr(n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))


def dst_expm_laplacian(dims, x, t):
    """exp(-t A) x for the Dirichlet Laplacian on `dims` (closed form)."""
    from scipy.fft import dstn, idstn
    dims = tuple(int(d) for d in dims)
    X = np.asarray(x, dtype=np.float64).reshape(dims[::-1])  # C order: last axis = first grid axis
    lam = 0.0
    grids = np.meshgrid(*[2.0 - 2.0 * np.cos(np.arange(1, d + 1) * np.pi / (d + 1)) for d in dims[::-1]],
                        indexing="ij")
    for g in grids:
        lam = lam + g
    Y = dstn(X, type=1, norm="ortho")
    Y *= np.exp(-t * lam)
    return idstn(Y, type=1, norm="ortho").reshape(-1)


def anisotropic_2d(m, seed, sigma=1.0, xscale=10.0):
    """2-D 5-point anisotropic diffusion operator on an m x m grid with
    Dirichlet ghosts.  Every node-to-neighbour edge (including the edges to
    the boundary ghosts) has a log-normal conductance exp(sigma * N(0,1));
    x-direction edges are multiplied by `xscale`.  A_ii = sum of the
    conductances of node i's 4 edges, A_ij = -c_ij.  SPD."""
    n = m * m
    # x-edges: (m+1) per grid row, y-edges: (m+1) per grid column
    nx = (m + 1) * m
    g = normal(seed, 2 * nx)
    cx = xscale * np.exp(sigma * g[:nx]).reshape(m, m + 1)   # cx[iy, k]: edge between ix=k-1 and ix=k
    cy = np.exp(sigma * g[nx:]).reshape(m, m + 1)            # cy[ix, k]: edge between iy=k-1 and iy=k
    iy, ix = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    ix = ix.reshape(-1)
    iy = iy.reshape(-1)
    idx = ix + m * iy
    diag = cx[iy, ix] + cx[iy, ix + 1] + cy[ix, iy] + cy[ix, iy + 1]
    rows = [idx]
    cols = [idx]
    vals = [diag]
    ok = ix > 0
    rows.append(idx[ok]); cols.append(idx[ok] - 1); vals.append(-cx[iy[ok], ix[ok]])
    ok = ix < m - 1
    rows.append(idx[ok]); cols.append(idx[ok] + 1); vals.append(-cx[iy[ok], ix[ok] + 1])
    ok = iy > 0
    rows.append(idx[ok]); cols.append(idx[ok] - m); vals.append(-cy[ix[ok], iy[ok]])
    ok = iy < m - 1
    rows.append(idx[ok]); cols.append(idx[ok] + m); vals.append(-cy[ix[ok], iy[ok] + 1])
    return _coo_to_csr(n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))


def rgg_laplacian(n, seed, degree=8.0, wlo=0.5, whi=1.5):
    """Random geometric graph in the unit square: n uniform points, an edge
    for every pair closer than r = sqrt(degree/(pi n)), weights U[wlo, whi].
    Returns the graph Laplacian L = D - W (PSD, one zero eigenvalue per
    connected component)."""
    from scipy.spatial import cKDTree
    pts = uniform01(seed, 2 * n).reshape(n, 2)
    r = np.sqrt(degree / (np.pi * n))
    pairs = cKDTree(pts).query_pairs(r, output_type="ndarray")
    pairs = pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))]
    ne = pairs.shape[0]
    w = wlo + (whi - wlo) * uniform01(seed ^ 0x5DEECE66D, ne)
    i, j = pairs[:, 0].astype(np.int64), pairs[:, 1].astype(np.int64)
    deg = np.bincount(i, weights=w, minlength=n) + np.bincount(j, weights=w, minlength=n)
    rows = np.concatenate([np.arange(n), i, j])
    cols = np.concatenate([np.arange(n), j, i])
    vals = np.concatenate([deg, -w, -w])
    return _coo_to_csr(n, rows, cols, vals)


def chung_lu_laplacian(n, seed, degree=8.0, exponent=2.2, wlo=0.5, whi=1.5):
    """Graph with a power-law degree profile (Chung-Lu: about n*degree/2
    edges whose ends are drawn with probability proportional to the weights
    w_i = (i + 1)^(-1/(exponent - 1)), repeated pairs merged, self loops
    dropped), edge weights U[wlo, whi], nodes numbered by decreasing weight.
    Returns the Laplacian L = D - W: rows of very uneven length (a few
    thousand entries at the top, 1-2 at the tail), the irregular case no
    sliced or diagonal storage serves."""
    w = (np.arange(n, dtype=np.float64) + 1.0) ** (-1.0 / (exponent - 1.0))
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    m = int(n * degree / 2)
    u = uniform01(seed, 2 * m)
    i = np.minimum(np.searchsorted(cdf, u[:m], side="right"), n - 1)
    j = np.minimum(np.searchsorted(cdf, u[m:], side="right"), n - 1)
    keep = i != j
    a, b = np.minimum(i[keep], j[keep]), np.maximum(i[keep], j[keep])
    key = np.unique(a.astype(np.int64) * n + b)
    a, b = key // n, key % n
    ww = wlo + (whi - wlo) * uniform01(seed ^ 0x2545F491, len(a))
    deg = np.bincount(a, weights=ww, minlength=n) + np.bincount(b, weights=ww, minlength=n)
    rows = np.concatenate([np.arange(n), a, b])
    cols = np.concatenate([np.arange(n), b, a])
    vals = np.concatenate([deg, -ww, -ww])
    return _coo_to_csr(n, rows, cols, vals)


def rgg_sizes(count, seed, nmin=1000, nmax=64000):
    """Graph orders drawn log-uniform in [nmin, nmax]."""
    u = uniform01(seed, count)
    return np.floor(np.exp(np.log(nmin) + u * (np.log(nmax) - np.log(nmin)))).astype(np.int64)


def to_scipy(csr, n=None):
    import scipy.sparse as sp
    rp, ci, va = csr
    n = len(rp) - 1 if n is None else n
    return sp.csr_matrix((va, ci, rp), shape=(n, n))
