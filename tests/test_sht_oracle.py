"""The CPU restatement of the spin-0 general synthesis (oracle/sht.py)
pinned independently: the Legendre recursion against scipy's spherical
harmonics (which carry the Condon-Shortley phase), the synthesis against
the full two-sided sum over m = -l..l, orthonormality under Gauss-Legendre
quadrature, and the adjoint identity of test_sht.py:238-255."""
import numpy as np
import pytest
from scipy.special import sph_harm_y

from oracle import sht as S


def _loc(rng, n):
    loc = rng.uniform(0., 1., (n, 2))
    loc[:, 0] *= np.pi
    loc[:, 1] *= 2 * np.pi
    return loc


@pytest.mark.parametrize("lmax", [0, 1, 5, 17, 40])
def test_lambda_matches_scipy(lmax):
    rng = np.random.default_rng(lmax)
    loc = _loc(rng, 50)
    # poles included
    loc[0] = (0.0, 1.0)
    loc[1] = (np.pi, 2.0)
    for m in range(lmax + 1):
        lam = S.lambda_m(lmax, m, loc[:, 0])
        for l in range(m, lmax + 1):
            y = sph_harm_y(l, m, loc[:, 0], loc[:, 1])
            np.testing.assert_allclose(lam[l - m] * np.exp(1j * m * loc[:, 1]),
                                       y, rtol=1e-11, atol=1e-13)


def test_synthesis_is_the_two_sided_sum():
    rng = np.random.default_rng(3)
    lmax = 12
    a = S.random_alm(lmax, lmax, rng)
    loc = _loc(rng, 40)
    ms = S.default_mstart(lmax)
    ref = np.zeros(len(loc), np.complex128)
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            am = a[ms[abs(m)] + l]
            alm = am if m >= 0 else (-1) ** m * np.conj(am)
            ref += alm * sph_harm_y(l, m, loc[:, 0], loc[:, 1])
    assert np.max(np.abs(ref.imag)) < 1e-12
    got = S.synthesis_general(a, lmax, loc)
    np.testing.assert_allclose(got, ref.real, rtol=1e-12, atol=1e-12)


def test_orthonormality_under_quadrature():
    lmax = 10
    x, w = np.polynomial.legendre.leggauss(lmax + 2)
    th = np.arccos(x)
    for m in range(lmax + 1):
        lam = S.lambda_m(lmax, m, th)
        # int Y_lm conj(Y_l'm) dOmega = 2 pi sum_j w_j lam_l lam_l'
        G = 2 * np.pi * (lam * w) @ lam.T
        np.testing.assert_allclose(G, np.eye(lmax + 1 - m), atol=1e-13)


@pytest.mark.parametrize("lmax,mmax,lstride", [(10, 8, 1), (16, 16, 1),
                                               (9, 9, 2)])
def test_adjoint_identity(lmax, mmax, lstride):
    rng = np.random.default_rng(48)
    ms = S.default_mstart(lmax, mmax)
    if lstride != 1:
        ms = ms * lstride
    n = int(max(ms[m] + lmax * lstride for m in range(mmax + 1))) + 1
    a = np.zeros(n, np.complex128)
    for m in range(mmax + 1):
        for l in range(m, lmax + 1):
            a[ms[m] + l * lstride] = rng.uniform(-1, 1) + \
                (1j * rng.uniform(-1, 1) if m else 0)
    loc = _loc(rng, 33)
    r = rng.uniform(-0.5, 0.5, len(loc))
    f = S.synthesis_general(a, lmax, loc, ms, lstride)
    b = S.adjoint_synthesis_general(r, lmax, loc, ms, lstride, size=n)
    v1 = S.almdot(a, b, lmax, ms, lstride)
    v2 = float(np.dot(r, f))
    assert abs(v1 - v2) <= 1e-12 * max(abs(v1), 1.0)


def test_a_planted_sign_error_is_caught():
    """flipping the Condon-Shortley phase changes odd-m values, which the
    scipy comparison detects"""
    lmax, m = 6, 3
    th = np.linspace(0.1, 3.0, 7)
    lam = -S.lambda_m(lmax, m, th)
    y = sph_harm_y(4, m, th, 0.0 * th)
    assert not np.allclose(lam[4 - m], y.real)


def test_matches_ducc0_itself():
    """the restatement against ducc0's own synthesis_general and
    adjoint_synthesis_general (compiled from the reference sources into
    oracle/_ref) at eps 1e-10, both precisions' layouts"""
    from oracle import ref as R
    from oracle import core as O
    if not R.available():
        pytest.skip("reference library not built")
    rng = np.random.default_rng(1)
    lmax = 40
    a = S.random_alm(lmax, lmax, rng)
    loc = _loc(rng, 500)
    f = R.sht_synthesis_general(a, lmax, loc, 1e-10)
    assert O.l2error(f.astype(complex), S.synthesis_general(a, lmax, loc)) \
        <= 1e-10
    r = rng.uniform(-1, 1, 500)
    b = R.sht_adjoint_synthesis_general(r, lmax, loc, 1e-10, len(a))
    assert O.l2error(b, S.adjoint_synthesis_general(r, lmax, loc)) <= 1e-10
