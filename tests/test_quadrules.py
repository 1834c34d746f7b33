"""Host-side quadrature rules and special functions of libpsfun_hip.so
(no device needed): the reference's control values (examples/utils/utiltest.f90)
and agreement with the CPU oracle."""
import numpy as np

from oracle import oracle as O
from psfun_amd import quadrature as Q


def test_horner_kat():
    assert Q.horner([132.0, 42.0, 14.0, 5.0, 2.0, 1.0, 0.0], 1.0) == 196.0


def test_ellipkkp_kat_and_oracle():
    K = Q.ellipkkp(0.5)
    assert abs(K[0] - 1.588191701877384) < 1e-14 and abs(K[1] - 2.978718299395645) < 1e-14
    for L in (0.01, 0.2, 1.0, 4.0, 10.5):
        assert np.allclose(Q.ellipkkp(L), O.ellipkkp(L), rtol=1e-15, atol=0)


def test_ellipj_kats():
    sn, cn, dn = Q.ellipj(1.0, 0.5)
    assert abs(sn - 0.838274911024583) < 1e-14 and abs(cn - 0.545247809300255) < 1e-14
    assert abs(dn - 0.984699634947678) < 1e-14
    sn, cn, dn = Q.ellipj(0.5j, 1.0)
    assert abs(sn - 0.521141424070547j) < 1e-14 and abs(cn - 1.127647278133671) < 1e-14
    assert abs(dn - 1.000253555731494) < 1e-14


def test_ellipj_matches_oracle_complex():
    for L in (0.05, 0.4, 1.5):
        for u in (0.2 + 0.1j, -1.0 + 0.7j, 0.9 + 1.9j):
            a = Q.ellipj(u, L)
            b = O.ellipj(u, L)
            assert all(abs(complex(p) - complex(q)) <= 1e-13 * (1 + abs(q)) for p, q in zip(a, b))


def test_hhtmethod1_matches_oracle_and_example():
    f = lambda z: np.sqrt(1 + z)
    xi, c, eta, sign, info = Q.hhtmethod1(f, 10, rparams=[0.01, 4.0])
    xo, co, eo, so = O.hhtmethod1(f, 10, [0.01, 4.0])
    assert info == 0 and sign == so == -1.0
    assert np.allclose(xi, xo, rtol=1e-14) and np.allclose(c, co, rtol=1e-13) and abs(eta - eo) < 1e-15
    # rparams order does not matter (min/max taken, quadrules:65-66)
    xi2, *_ = Q.hhtmethod1(f, 10, rparams=[4.0, 0.01])
    assert np.array_equal(xi, xi2)


def test_talbot_matches_oracle():
    xi, c, eta, sign, info = Q.talbot(None, 16, rparams=[0.5])
    so, wo = O.talbot(16, 0.5)
    assert info == 0 and eta == 1.0 and sign == 1.0
    assert np.allclose(xi, so, rtol=1e-14) and np.allclose(c, wo, rtol=1e-13)


def test_fold_conjugate_pairs():
    xi, c, *_ = Q.talbot(None, 16, rparams=[1.0])
    s, w = Q.fold_conjugate_pairs(xi, c)
    assert len(s) == 8
    for lam in (0.0, 0.3, 5.0, 11.9):
        full = np.real(np.sum(c / (xi + lam)))
        half = np.real(np.sum(w / (s + lam)))
        assert abs(full - half) < 1e-14 * (1 + abs(full))


def test_computepoles_prefers_cparams():
    # psfun_z_quadrature_mod.F90:130-138: cparams is tested first
    seen = {}

    def rule(zfun, N, cparams=None, rparams=None):
        seen["c"], seen["r"] = cparams, rparams
        return np.ones(N, complex), np.ones(N, complex), 1.0, -1.0, 0

    q = Q.psfun_z_quadrature()
    assert q.computepoles(rule, None, 3, cparams=[1j], rparams=[1.0, 2.0]) == 0
    assert seen["c"] == [1j] and seen["r"] is None
