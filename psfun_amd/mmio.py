"""MatrixMarket ingest for the PSFUN drivers' matrices and vectors.

The reference drivers read their matrix with PSBLAS `mm_mat_read(a, info,
iunit, filename)` and an optional right-hand side with `mm_array_read(x,
info, iunit, filename)` (examples/polynomial/lanczostest.F90:107-117, the
same in arnoldi/serial/shifted drivers); PSBLAS itself is not in the
reference tree, so the format is taken from the MatrixMarket definition
(NIST, "The Matrix Market Exchange Formats: Initial Design", 1996):

  %%MatrixMarket matrix coordinate <real|integer|pattern> <general|symmetric|skew-symmetric>
  % comments
  M N NNZ
  i j [value]          (1-based, NNZ lines)

  %%MatrixMarket matrix array <real|integer> general
  M N
  value                (column-major, M*N lines)

Symmetric and skew-symmetric files store one triangle; the other is
generated (a_ji = a_ij, resp. -a_ij; diagonal entries once).  Repeated
entries are summed, as PSBLAS assembly sums them.  `pattern` entries get the
value 1.  Complex files and Harwell-Boeing (`hb_read`) are not read:
info = psb_err_from_subroutine_ with a message on stderr.

mm_mat_read(filename, ctxt=None) -> (a, info)    a: psb.psb_dspmat_type or None
mm_mat_read_csr(filename)        -> (rowptr, colind, val, nrows, ncols)   (host, 0-based)
mm_array_read(filename)          -> (x, info)    x: numpy float64 (M,) or (M, N)
mm_array_write(x, title, filename) -> info
"""
import sys

import numpy as np

from . import psb


class MMError(ValueError):
    pass


def _header(line):
    t = line.strip().split()
    if len(t) < 2 or t[0].lower() != "%%matrixmarket" or t[1].lower() != "matrix":
        raise MMError(f"not a MatrixMarket matrix header: {line.strip()!r}")
    if len(t) != 5:
        raise MMError(f"MatrixMarket header needs format, field and symmetry: {line.strip()!r}")
    return t[2].lower(), t[3].lower(), t[4].lower()


def _data_lines(fh):
    for line in fh:
        s = line.strip()
        if s and not s.startswith("%"):
            yield s


def mm_mat_read_csr(filename):
    """Host CSR (0-based int32 rowptr/colind, float64 val) of a coordinate file."""
    with open(filename, "r") as fh:
        fmt, field, sym = _header(fh.readline())
        if fmt != "coordinate":
            raise MMError(f"mm_mat_read needs a coordinate file, not {fmt!r}")
        if field not in ("real", "integer", "double", "pattern"):
            raise MMError(f"unsupported field {field!r}")
        if sym not in ("general", "symmetric", "skew-symmetric"):
            raise MMError(f"unsupported symmetry {sym!r}")
        lines = _data_lines(fh)
        try:
            m, n, nz = (int(v) for v in next(lines).split()[:3])
        except (StopIteration, ValueError):
            raise MMError("missing or malformed size line")
        ncol = 2 if field == "pattern" else 3
        body = list(lines)
    if len(body) < nz:
        raise MMError(f"expected {nz} entries, found {len(body)}")
    if nz:
        arr = np.array([ln.split()[:ncol] for ln in body[:nz]], dtype=np.float64)
        if arr.shape[1] != ncol:
            raise MMError("malformed entry line")
        ii = arr[:, 0].astype(np.int64) - 1
        jj = arr[:, 1].astype(np.int64) - 1
        vv = np.ones(nz) if field == "pattern" else arr[:, 2].copy()
    else:
        ii = jj = np.zeros(0, np.int64)
        vv = np.zeros(0)
    if nz and (ii.min() < 0 or jj.min() < 0 or ii.max() >= m or jj.max() >= n):
        raise MMError("entry index outside the declared size")
    if sym != "general":
        off = ii != jj
        sign = -1.0 if sym == "skew-symmetric" else 1.0
        ii, jj, vv = (np.concatenate([ii, jj[off]]), np.concatenate([jj, ii[off]]),
                      np.concatenate([vv, sign * vv[off]]))
    # CSR with repeated entries summed, columns ascending inside each row
    order = np.lexsort((jj, ii))
    ii, jj, vv = ii[order], jj[order], vv[order]
    if len(ii):
        keep = np.ones(len(ii), bool)
        keep[1:] = (ii[1:] != ii[:-1]) | (jj[1:] != jj[:-1])
        grp = np.cumsum(keep) - 1
        vs = np.zeros(int(keep.sum()))
        np.add.at(vs, grp, vv)
        ii, jj, vv = ii[keep], jj[keep], vs
    rowptr = np.zeros(m + 1, np.int64)
    np.add.at(rowptr, ii + 1, 1)
    rowptr = np.cumsum(rowptr)
    return rowptr.astype(np.int32), jj.astype(np.int32), vv.astype(np.float64), m, n


def mm_mat_read(filename, ctxt=None):
    """psb_dspmat_type from a MatrixMarket coordinate file; (a, info)."""
    try:
        rp, ci, va, m, n = mm_mat_read_csr(filename)
    except (OSError, MMError) as e:
        sys.stderr.write(f"mm_mat_read: {e}\n")
        return None, psb.psb_err_from_subroutine_
    if m != n:
        sys.stderr.write(f"mm_mat_read: the matrix is {m} x {n}; the f(A)x path needs a square matrix\n")
        return None, psb.psb_err_from_subroutine_
    return psb.psb_dspmat_type(rp, ci, va, 0, ctxt), psb.psb_success_


def mm_array_read(filename):
    """Dense array from a MatrixMarket array file; (x, info), x of shape (M,)
    for a single column, else (M, N) column-major as stored."""
    try:
        with open(filename, "r") as fh:
            fmt, field, sym = _header(fh.readline())
            if fmt != "array" or field not in ("real", "integer", "double") or sym != "general":
                raise MMError(f"mm_array_read needs 'array real general', not {fmt} {field} {sym}")
            lines = _data_lines(fh)
            m, n = (int(v) for v in next(lines).split()[:2])
            vals = np.array([float(ln.split()[0]) for ln in lines], dtype=np.float64)
        if len(vals) < m * n:
            raise MMError(f"expected {m * n} values, found {len(vals)}")
    except (OSError, MMError, StopIteration, ValueError) as e:
        sys.stderr.write(f"mm_array_read: {e}\n")
        return None, psb.psb_err_from_subroutine_
    x = vals[: m * n].reshape((n, m)).T
    return (x[:, 0].copy() if n == 1 else np.ascontiguousarray(x)), psb.psb_success_


def mm_array_write(x, title, filename):
    """Dense vector or column-major matrix as a MatrixMarket array file (the
    serial driver's mm_array_write(y, title, info, iunit, filename)); returns
    info."""
    x = np.asarray(x, dtype=np.float64)
    X = x.reshape(-1, 1) if x.ndim == 1 else x
    try:
        with open(filename, "w") as fh:
            fh.write("%%MatrixMarket matrix array real general\n")
            fh.write(f"% {title}\n")
            fh.write(f"{X.shape[0]} {X.shape[1]}\n")
            for v in X.T.reshape(-1):
                fh.write(f"{v:.17e}\n")
    except OSError as e:
        sys.stderr.write(f"mm_array_write: {e}\n")
        return psb.psb_err_from_subroutine_
    return psb.psb_success_


def hb_read(filename, ctxt=None):
    """Harwell-Boeing input is not read (the reference's HB branch needs
    PSBLAS's hb_read); returns (None, psb_err_from_subroutine_)."""
    sys.stderr.write("hb_read: Harwell-Boeing input is not supported; convert the matrix to MatrixMarket\n")
    return None, psb.psb_err_from_subroutine_
