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
value 1.  Complex files are not read: info = psb_err_from_subroutine_ with a
message on stderr.

Harwell-Boeing files (the drivers' filefmt 'HB', hb_read(a, info, iunit, b,
filename), lanczostest.F90:119-122; format from Duff, Grimes and Lewis,
"Users' Guide for the Harwell-Boeing Sparse Matrix Collection", 1992):

  line 1  title (A72), key (A8)
  line 2  TOTCRD PTRCRD INDCRD VALCRD RHSCRD          (5I14)
  line 3  MXTYPE (A3), 11 blanks, NROW NCOL NNZERO NELTVL (4I14)
  line 4  PTRFMT (A16) INDFMT (A16) VALFMT (A20) RHSFMT (A20)
  line 5  RHSTYP (A3), 11 blanks, NRHS NRHSIX (2I14)   only when RHSCRD > 0
  then NCOL+1 column pointers, NNZERO row indices (1-based, compressed
  columns) and NNZERO values, each block on its own cards in its Fortran
  format, then the right-hand sides (type 'F': NRHS full columns).

MXTYPE R (real) or P (pattern: values 1), U (unsymmetric), S (symmetric) or
Z (skew-symmetric: one triangle stored), A (assembled).  The data cards are
read as fixed-width fields of the widths their formats give, so values that
touch (no blank between them) and D exponents read as Fortran reads them.

mm_mat_read(filename, ctxt=None) -> (a, info)    a: psb.psb_dspmat_type or None
mm_mat_read_csr(filename)        -> (rowptr, colind, val, nrows, ncols)   (host, 0-based)
mm_array_read(filename)          -> (x, info)    x: numpy float64 (M,) or (M, N)
mm_array_write(x, title, filename) -> info
hb_read(filename, ctxt=None)     -> (a, b, info)  b: numpy (M, NRHS) or None
hb_read_csr(filename)            -> dict(rowptr, colind, val, m, n, rhs, title, key, mxtype)
hb_write(filename, rowptr, colind, val, title, key, mxtype, rhs=None) -> info
"""
import re
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


_HB_FMT = re.compile(r"(\d*)\s*\(?\s*([IEDFG])\s*(\d+)", re.I)


def _hb_format(fmt):
    """(fields per card, field width) of a Fortran data format such as
    (10I8), (4E20.12), (1P,5D16.8), (1P4E20.12) or (3(1PE25.16))."""
    f = re.sub(r"\d*\s*P\s*,?", "", fmt.strip().upper().replace(" ", ""), count=0)
    m = _HB_FMT.search(f.strip("()"))
    if not m:
        raise MMError(f"unsupported Harwell-Boeing format {fmt!r}")
    return int(m.group(1) or 1), int(m.group(3))


def _hb_float(t):
    t = t.strip().upper().replace("D", "E")
    if not t:
        raise MMError("empty Harwell-Boeing value field")
    # Fortran accepts an exponent without its letter: 1.5-03, 2.0+10
    t = re.sub(r"(?<=[0-9.])([+-])(?=\d+$)", r"E\1", t) if "E" not in t else t
    return float(t)


def _hb_block(lines, start, ncards, fmt, count, conv):
    per, width = _hb_format(fmt)
    out = []
    for ln in lines[start:start + ncards]:
        ln = ln.rstrip("\n")
        for k in range(per):
            field = ln[k * width:(k + 1) * width]
            if not field.strip():
                continue
            out.append(conv(field))
            if len(out) == count:
                return out
    if len(out) < count:
        raise MMError(f"expected {count} entries, found {len(out)}")
    return out


def hb_read_csr(filename):
    """Host CSR (0-based) of a Harwell-Boeing file, its right-hand sides and
    header (see the module docstring for the layout)."""
    with open(filename, "r") as fh:
        lines = fh.readlines()
    if len(lines) < 4:
        raise MMError("Harwell-Boeing header needs 4 lines")
    title, key = lines[0][:72].rstrip(), lines[0][72:80].strip()

    def ints(line, k):
        t = line.split()
        if len(t) < k:
            raise MMError(f"bad Harwell-Boeing header line {line.rstrip()!r}")
        return [int(v) for v in t[:k]]

    t2 = lines[1].split()
    totcrd, ptrcrd, indcrd, valcrd = (int(v) for v in t2[:4])
    rhscrd = int(t2[4]) if len(t2) > 4 else 0
    mxtype = lines[2][:3].upper()
    nrow, ncol, nnz = ints(lines[2][3:], 3)
    f4 = lines[3].rstrip("\n")
    ptrfmt, indfmt, valfmt, rhsfmt = f4[0:16], f4[16:32], f4[32:52], f4[52:72]
    if mxtype[0] not in "RP" or mxtype[1] not in "USZR" or mxtype[2] != "A":
        raise MMError(f"Harwell-Boeing type {mxtype!r}: real or pattern, assembled, is read")
    start = 4
    nrhs, rhstyp = 0, ""
    if rhscrd > 0:
        rhstyp = lines[4][:3].upper()
        nrhs = ints(lines[4][3:], 1)[0]
        start = 5
    colptr = np.array(_hb_block(lines, start, ptrcrd, ptrfmt, ncol + 1, lambda f: int(f)), np.int64)
    rowind = np.array(_hb_block(lines, start + ptrcrd, indcrd, indfmt, nnz, lambda f: int(f)), np.int64)
    if mxtype[0] == "P":
        vals = np.ones(nnz)
    else:
        vals = np.array(_hb_block(lines, start + ptrcrd + indcrd, valcrd, valfmt, nnz, _hb_float))
    rhs = None
    if nrhs > 0 and rhstyp[:1] == "F":
        r = _hb_block(lines, start + ptrcrd + indcrd + valcrd, rhscrd, rhsfmt or valfmt, nrow * nrhs, _hb_float)
        rhs = np.array(r).reshape(nrhs, nrow).T.copy()
    cols = np.repeat(np.arange(ncol), np.diff(colptr))
    rows = rowind - 1
    if mxtype[1] in "SZ":
        off = rows != cols
        sgn = -1.0 if mxtype[1] == "Z" else 1.0
        rows, cols, vals = (np.concatenate([rows, cols[off]]), np.concatenate([cols, rows[off]]),
                            np.concatenate([vals, sgn * vals[off]]))
    order = np.lexsort((cols, rows))          # row-major; a stable sort keeps repeats in file order
    rows, cols, vals = rows[order], cols[order], vals[order]
    rowptr = np.zeros(nrow + 1, np.int64)
    np.add.at(rowptr, rows + 1, 1)
    rowptr = np.cumsum(rowptr)
    return dict(rowptr=rowptr.astype(np.int32), colind=cols.astype(np.int32), val=vals.astype(np.float64),
                m=nrow, n=ncol, rhs=rhs, title=title, key=key, mxtype=mxtype, totcrd=totcrd)


def hb_read(filename, ctxt=None):
    """psb_dspmat_type from a Harwell-Boeing file and its first right-hand
    sides when it holds any; (a, b, info)."""
    try:
        h = hb_read_csr(filename)
    except (OSError, MMError, ValueError) as e:
        sys.stderr.write(f"hb_read: {e}\n")
        return None, None, psb.psb_err_from_subroutine_
    if h["m"] != h["n"]:
        sys.stderr.write(f"hb_read: the matrix is {h['m']} x {h['n']}; the f(A)x path needs a square matrix\n")
        return None, None, psb.psb_err_from_subroutine_
    a = psb.psb_dspmat_type(h["rowptr"], h["colind"], h["val"], 0, ctxt)
    return a, h["rhs"], psb.psb_success_


def _hb_cards(values, per, width, fmt_one):
    lines = []
    for k in range(0, len(values), per):
        lines.append("".join(fmt_one(v).rjust(width) for v in values[k:k + per]))
    return lines


def hb_write(filename, rowptr, colind, val, title, key, mxtype="RUA", rhs=None):
    """Write a square CSR matrix (0-based) as a Harwell-Boeing file: RUA
    stores every entry by columns, RSA the lower triangle (the matrix must
    be symmetric); values in E25.16 (exact in double precision).  Returns
    info."""
    rp, ci, va = (np.asarray(rowptr), np.asarray(colind), np.asarray(val, np.float64))
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    cols = ci.astype(np.int64)
    if mxtype.upper() == "RSA":
        keep = rows >= cols
        rows, cols, va = rows[keep], cols[keep], va[keep]
    order = np.lexsort((rows, cols))          # column-major
    rows, cols, va = rows[order], cols[order], va[order]
    colptr = np.concatenate([[0], np.cumsum(np.bincount(cols, minlength=n))]) + 1
    nnz = len(va)
    iw = len(str(max(int(colptr[-1]), n))) + 1
    iper = max(1, 80 // iw)
    ptr = _hb_cards(list(colptr), iper, iw, str)
    ind = _hb_cards(list(rows + 1), iper, iw, str)
    vls = _hb_cards(list(va), 3, 25, lambda v: f"{v:.16E}")
    rl = []
    if rhs is not None:
        R = np.asarray(rhs, np.float64).reshape(n, -1)
        rl = _hb_cards(list(R.T.reshape(-1)), 3, 25, lambda v: f"{v:.16E}")
    try:
        with open(filename, "w") as fh:
            fh.write(f"{title[:72]:<72}{key[:8]:<8}\n")
            fh.write(f"{len(ptr) + len(ind) + len(vls) + len(rl):14d}{len(ptr):14d}{len(ind):14d}{len(vls):14d}"
                     f"{len(rl):14d}\n")
            fh.write(f"{mxtype.upper():<3}{'':11}{n:14d}{n:14d}{nnz:14d}{0:14d}\n")
            fmt_i = f"({iper}I{iw})"
            fh.write(f"{fmt_i:<16}{fmt_i:<16}{'(3E25.16)':<20}{'(3E25.16)' if rl else '':<20}\n")
            if rl:
                fh.write(f"{'F':<3}{'':11}{R.shape[1]:14d}{0:14d}\n")
            for ln in ptr + ind + vls + rl:
                fh.write(ln + "\n")
    except OSError as e:
        sys.stderr.write(f"hb_write: {e}\n")
        return psb.psb_err_from_subroutine_
    return psb.psb_success_
