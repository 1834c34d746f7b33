"""The PSFUN example drivers on the device library, run from the same input
decks (examples/polynomial/{lanczos,arnoldi}.inp, examples/serial/serial.inp,
examples/shiftedkrylov/shifted.inp):

  python -m psfun_amd.drivers lanczos  [deck]     (lanczostest.F90)
  python -m psfun_amd.drivers arnoldi  [deck]     (arnolditest.F90)
  python -m psfun_amd.drivers serial   [deck]     (serialtest.f90)
  python -m psfun_amd.drivers shifted  [deck]     (shiftedtest.f90)
  python -m psfun_amd.drivers quadrature           (quadraturetest.F90: the poles)

As the Fortran programs do, the deck comes from the file named as the first
argument or else from standard input: a count line, then one value per line
(Fortran list-directed: the first item of the line, quotes and 'd'
exponents allowed, T/F logicals).  Matrices are MatrixMarket (MM) or
Harwell-Boeing (HB) files, vectors MatrixMarket arrays (an HB file's own
right-hand side, when it holds one, is the vector); the partitioning and storage-format
entries are read and reported but a single device holds the whole matrix.

Kept from the reference drivers: the Lanczos and Arnoldi programs run with
eps = 1e-6, itmax = 100, itrace = 1, istop = 1 whatever the deck holds
(lanczostest.F90:213-216, arnolditest.F90:218-221); x = (1,...,1) when the
deck names NONE; the shifted program's printed residual is ||b - A x|| with
the unshifted A (shiftedtest.f90:248-251), printed here beside the residual
of the shifted system.
"""
import os
import sys
import time

import numpy as np
import torch

from . import mmio, psb
from .krylov import psfun_d_krylov
from .serial import psfun_d_serial
from .shifted import psb_dprec_type, psb_krylov


class DeckError(ValueError):
    pass


def _item(line):
    s = line.strip()
    if not s:
        raise DeckError("empty deck line")
    if s[0] in "'\"":
        q = s[0]
        end = s.find(q, 1)
        if end < 0:
            raise DeckError(f"unterminated string: {line!r}")
        return s[1:end]
    return s.replace(",", " ").split()[0]


def read_deck(src):
    """Values of a deck: `src` is a path, a file object or the deck text."""
    if hasattr(src, "read"):
        text = src.read()
    elif isinstance(src, str) and "\n" not in src and os.path.exists(src):
        with open(src) as fh:
            text = fh.read()
    else:
        text = str(src)
    lines = [ln for ln in text.splitlines() if ln.strip()]
    if not lines:
        raise DeckError("empty deck")
    n = int(_item(lines[0]))
    if len(lines) - 1 < n:
        raise DeckError(f"deck announces {n} values, holds {len(lines) - 1}")
    return [_item(ln) for ln in lines[1:n + 1]]


def _real(v):
    return float(str(v).lower().replace("d", "e"))


def _logical(v):
    t = str(v).strip().lower().strip(".")
    if t in ("t", "true"):
        return True
    if t in ("f", "false"):
        return False
    raise DeckError(f"not a logical: {v!r}")


def _vec(a):
    return psb.psb_d_vect_type(torch.as_tensor(np.ascontiguousarray(a, np.float64), device="cuda:0"))


def _read_matrix(name, filefmt, ctxt):
    """(matrix, right-hand sides held in the file or None): MM reads the
    matrix file alone, HB the matrix and its right-hand sides
    (lanczostest.F90:107-127)."""
    f = str(filefmt).upper()
    if f == "MM":
        a, info = mmio.mm_mat_read(name, ctxt)
        b = None
    elif f == "HB":
        a, b, info = mmio.hb_read(name, ctxt)
    else:
        raise DeckError(f"file format {filefmt!r}: MM or HB")
    if info != 0:
        raise DeckError(f"could not read the matrix {name!r}")
    return a, b


def _rhs(name, n, file_rhs=None, filefmt="MM"):
    """The first right-hand side of an HB file when it holds one of length n
    (the deck's vector name is not read then), the deck's MM vector file, or
    (1, ..., 1) (lanczostest.F90:136-155)."""
    if str(filefmt).upper() == "HB":
        if file_rhs is not None and file_rhs.shape[0] == n:
            return file_rhs[:, 0].copy()
        return np.ones(n)
    if str(name).upper() == "NONE":
        return np.ones(n)
    x, info = mmio.mm_array_read(name)
    if info != 0 or x.shape[0] != n:
        raise DeckError(f"could not read a vector of length {n} from {name!r}")
    return x if x.ndim == 1 else x[:, 0]


def krylov_driver(kname, deck, out=sys.stdout):
    """lanczostest.F90 / arnolditest.F90."""
    v = read_deck(deck)
    if len(v) not in (8, 12):
        raise DeckError(f"the {kname.lower()} deck holds 8 or 12 values, not {len(v)}")
    mname, rhs_file, filefmt, part, afmt, fname, variant = v[:7]
    scaling = _real(v[7])
    ctx = psb.default_context()
    t0 = time.perf_counter()
    a, frhs = _read_matrix(mname, filefmt, ctx)
    n = a.get_nrows()
    x = _vec(_rhs(rhs_file, n, frhs, filefmt))
    y = _vec(np.zeros(n))
    t_read = time.perf_counter() - t0
    fun = psfun_d_serial()
    fun.set("FNAME", fname)
    fun.set("VARIANT", variant)
    fun.set("SCALING", scaling)
    km = psfun_d_krylov()
    km.set("KNAME", kname)
    eps, itmax, itrace, istop = 1e-6, 100, 1, 1     # the drivers' fixed values
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    info, it, err, res = km.apply(fun, a, psb.psb_desc_type(n, ctx), y, x, eps, itmax=itmax, itrace=itrace,
                                  istop=istop)
    torch.cuda.synchronize()
    t2 = time.perf_counter() - t1
    w = out.write
    w(f"Solving matrix       : {mname}\nPartitioning         : {part}\nStorage format       : {afmt}\n"
      f"Function to compute  : {fname}\nAlgorithmic variant  : {variant}\nScaling              : {scaling:8.2f}\n")
    w(f"Time to read and partition matrix : {t_read:12.5e}\n")
    w(f"Matrix size                   : {n:12d}\nTime to solve system          : {t2:12.5e}\n"
      f"Time per iteration            : {t2 / max(it, 1):12.5e}\nNumber of iterations          : {it:12d}\n"
      f"Convergence indicator on exit : {err:12.5e}\nInfo  on exit                 : {info:12d}\n")
    return {"info": info, "iter": it, "err": err, "res": res, "y": y.get_vect(), "time": t2}


def serial_driver(deck, out=sys.stdout):
    """serialtest.f90: y = f(scaling*A) 1 on the whole matrix; with dump,
    the result goes to result.mtx."""
    v = read_deck(deck)
    if len(v) < 5:
        raise DeckError(f"the serial deck holds 5 values, not {len(v)}")
    mname, fname, variant = v[:3]
    scaling = _real(v[3])
    dump = _logical(v[4])
    rp, ci, va, m, n = mmio.mm_mat_read_csr(mname)
    fun = psfun_d_serial()
    fun.set("FNAME", fname)
    fun.set("VARIANT", variant)
    fun.set("SCALING", scaling)
    x = np.ones(n)
    y = np.zeros(n)
    t1 = time.perf_counter()
    info = fun.apply((rp, ci, va), y, x)
    t2 = time.perf_counter() - t1
    w = out.write
    w(f"Solving matrix       : {mname}\nFunction to compute  : {fname}\nAlgorithmic variant  : {variant}\n"
      f"Scaling              : {scaling:8.2f}\nMatrix of size matrix {n:6d}x{n:6d} nnz ={len(va):6d}\n"
      f"Elapsed time: {t2:12.5e}\n")
    if dump:
        mmio.mm_array_write(y, "serialtest out vector", "result.mtx")
    return {"info": info, "y": y, "time": t2}


def shifted_driver(deck, out=sys.stdout):
    """shiftedtest.f90: (eta A + zeta I) x = b by the shifted PCG."""
    v = read_deck(deck)
    if len(v) != 14:
        raise DeckError(f"the shifted deck holds 14 values, not {len(v)}")
    mname, bname = v[0], v[1]
    eta, zeta = _real(v[2]), _real(v[3])
    filefmt, kmethd, ptype, afmt, part = v[4:9]
    istopc, itmax, itrace, irst = (int(t) for t in v[9:13])
    eps = _real(v[13])
    ctx = psb.default_context()
    a, frhs = _read_matrix(mname, filefmt, ctx)
    n = a.get_nrows()
    bh = _rhs(bname, n, frhs, filefmt)
    b = _vec(bh)
    x = _vec(np.zeros(n))
    desc = psb.psb_desc_type(n, ctx)
    prec = psb_dprec_type()
    if prec.init(ctx, ptype) != 0:
        raise DeckError(f"preconditioner {ptype!r}: NONE, DIAG and BJAC are available")
    t0 = time.perf_counter()
    prec.build(a, desc)
    tprec = time.perf_counter() - t0
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    info, it, err, cond = psb_krylov(kmethd, a, prec, b, eta, zeta, x, eps, desc, itmax=itmax, itrace=itrace,
                                     irst=irst, istop=istopc, cond=True)
    torch.cuda.synchronize()
    t2 = time.perf_counter() - t1
    # the reference's check r = b - A x with the unshifted A (psb_spmm(-1,a,x,1,r)),
    # and the shifted system's residual b - (eta A + zeta I) x
    r_ref = _vec(bh)
    psb.psb_spmm(-1.0, a, x, 1.0, r_ref, desc)
    r_sys = _vec(bh)
    psb.psb_spmm(-eta, a, x, 1.0, r_sys, desc)
    psb.psb_geaxpby(-zeta, x, 1.0, r_sys, desc)
    resmx, _ = psb.psb_genrm2(r_ref, desc)
    resmxp, _ = psb.psb_geamax(r_ref, desc)
    ressys, _ = psb.psb_genrm2(r_sys, desc)
    xh = x.get_vect()
    w = out.write
    w(f"Matrix: {mname}\nη     : {eta:12.5e}\nζ     : {zeta:12.5e}\n"
      f"Iterations to convergence: {it:6d}\nError estimate on exit   : {err:12.5e}\n"
      f"Time to buil prec.       : {tprec:12.5e}\nTime to solve system     : {t2:12.5e}\n"
      f"Time per iteration       : {t2 / max(it, 1):12.5e}\nTotal time               : {t2 + tprec:12.5e}\n"
      f"Residual norm 2          : {resmx:12.5e}\n"
      f"Residual norm inf        : {resmxp:12.5e}\n"
      f"Shifted-system residual  : {ressys:12.5e}\n")
    return {"info": info, "iter": it, "err": err, "cond": cond, "x": xh, "time": t2,
            "residual_unshifted": resmx, "residual_shifted": ressys}


def quadrature_driver(out=sys.stdout):
    """quadraturetest.F90: the N = 10 poles of HHT method 1 for
    f(z) = sqrt(1 + z) on [0.01, 4] (lines 60-67), listed (the reference
    plots them through gnuplot)."""
    from . import quadrature as Q
    q = Q.psfun_z_quadrature()
    info = q.computepoles(Q.hhtmethod1, lambda z: np.sqrt(1.0 + z), 10, rparams=[0.01, 4.0])
    w = out.write
    w(f"HHT method 1, N = 10, [0.01, 4], f(z) = sqrt(1+z): eta = {q.eta:.16e}, sign = {q.sign:+.1f}\n")
    w(f"{'j':>3s} {'Re xi':>24s} {'Im xi':>24s} {'Re c':>24s} {'Im c':>24s}\n")
    for j, (xi, c) in enumerate(zip(q.xi, q.c), 1):
        w(f"{j:3d} {xi.real:24.16e} {xi.imag:24.16e} {c.real:24.16e} {c.imag:24.16e}\n")
    return {"info": info, "xi": np.asarray(q.xi), "c": np.asarray(q.c), "eta": q.eta, "sign": q.sign}


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    if not argv or argv[0] not in ("lanczos", "arnoldi", "serial", "shifted", "quadrature"):
        sys.stderr.write("usage: python -m psfun_amd.drivers {lanczos|arnoldi|serial|shifted} [deck]\n"
                         "       python -m psfun_amd.drivers quadrature\n")
        return 2
    if argv[0] == "quadrature":
        return 0 if quadrature_driver()["info"] == 0 else 1
    deck = argv[1] if len(argv) > 1 else sys.stdin
    try:
        if argv[0] in ("lanczos", "arnoldi"):
            r = krylov_driver(argv[0].upper(), deck)
        elif argv[0] == "serial":
            r = serial_driver(deck)
        else:
            r = shifted_driver(deck)
    except DeckError as e:
        sys.stderr.write(f"{argv[0]}: {e}\n")
        return 1
    return 0 if r["info"] == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
