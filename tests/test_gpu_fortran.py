"""The iso_c_binding boundary from Fortran: tests/fortran/psfun_driver (built
by tests/fortran/build.py against include/psfun_hip_mod.f90) runs the
reference example decks lanczos.inp / arnoldi.inp / shifted.inp on the 50x50
example matrix, the Lanczos deck with a Fortran USERF function, and the
batched exp(-tB)X with its one-call vector-Jacobian product; its results are
checked against the oracle, the fixture and dense scipy (expm, expm_frechet)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import REPO, load_golden
from oracle import oracle as O

pytestmark = pytest.mark.gpu

EXE = os.path.join(REPO, "tests", "fortran", "_build", "psfun_driver")


def _parse(text):
    out, cur = {}, None
    for line in text.strip().splitlines():
        parts = line.split()
        if parts[0].isalpha():
            cur = parts[0]
            out[cur] = dict(iter=int(parts[1]), err=float(parts[2]), y=[])
        else:
            out[cur]["y"].append(float(parts[0]))
    return out


def test_fortran_driver_matches_oracle(gpu_ctx):
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built (flang unavailable when building)")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = _parse(r.stdout)
    g = load_golden("laplacian50.npz")
    A = (g["rowptr"] - 1, g["colind"] - 1, g["val"])
    ones = np.ones(50)
    ref = O.lanczos(A, ones, 1e-6, 100, 1)
    y = np.array(got["LANCZOS"]["y"])
    assert got["LANCZOS"]["iter"] == ref["iter"]
    assert np.linalg.norm(y - ref["y"]) <= 1e-10 * np.linalg.norm(ref["y"])
    ref = O.arnoldi(A, ones, 1e-6, 100, 1)
    y = np.array(got["ARNOLDI"]["y"])
    assert got["ARNOLDI"]["iter"] == ref["iter"]
    assert np.linalg.norm(y - ref["y"]) <= 1e-10 * np.linalg.norm(ref["y"])
    ref = O.lanczos(A, ones, 1e-6, 100, 1, f=np.exp)      # USERF/SYM with the Fortran f(x) = exp(x)
    y = np.array(got["USERF"]["y"])
    assert got["USERF"]["iter"] == ref["iter"] == got["LANCZOS"]["iter"]
    assert np.linalg.norm(y - ref["y"]) <= 1e-10 * np.linalg.norm(ref["y"])
    x = np.array(got["CG"]["y"])
    assert np.linalg.norm(x - g["x_shift"]) <= 1e-6 * np.linalg.norm(g["x_shift"])
    assert got["CG"]["err"] <= 1e-7
    # batched exp(-tB)X and its vector-Jacobian product for B = -A (SPD),
    # against dense scipy: expm, and expm_frechet per stored nonzero
    import scipy.linalg as sla
    n, t = 50, 0.1
    rp, ci, va = A
    B = np.zeros((n, n))
    for i in range(n):
        for q in range(rp[i], rp[i + 1]):
            B[i, ci[q]] += -va[q]
    i1 = np.arange(1, n + 1)
    X = np.stack([np.ones(n), i1 / n], 1)
    G = np.stack([np.sin(0.3 * i1), np.cos(0.7 * i1)], 1)
    E = sla.expm(-t * B)
    Y = E @ X
    assert np.linalg.norm(np.array(got["BATCH"]["y"]) - Y.reshape(-1)) <= 1e-10 * np.linalg.norm(Y)
    gX = E @ G
    assert np.linalg.norm(np.array(got["VJPGX"]["y"]) - gX.reshape(-1)) <= 1e-10 * np.linalg.norm(gX)
    gt = -np.sum(G * (B @ Y))
    assert abs(got["VJPGX"]["err"] - gt) <= 1e-10 * abs(gt)
    gv = np.empty(len(va))
    for i in range(n):
        for q in range(rp[i], rp[i + 1]):
            Eq = np.zeros((n, n))
            Eq[i, ci[q]] = 1.0
            gv[q] = np.sum(G * (sla.expm_frechet(-t * B, -t * Eq, compute_expm=False) @ X))
    assert np.linalg.norm(np.array(got["VJPGV"]["y"]) - gv) <= 1e-8 * np.linalg.norm(gv)
