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


# ------------------------------------------------------------- spin > 0

def _goldberg(s, l, m, th):
    """lambda^s_lm(theta) from Goldberg et al.'s explicit sum for the
    spin-weighted harmonics (J. Math. Phys. 8, 2155 (1967), eq. 3.1), an
    independent closed form (factorials, no recursion)"""
    from math import comb, factorial
    if l < abs(s) or l < abs(m):
        return np.zeros_like(th)
    pre = (-1) ** m * np.sqrt(factorial(l + m) * factorial(l - m) * (2 * l + 1)
                              / (4 * np.pi * factorial(l + s) * factorial(l - s)))
    tot = np.zeros_like(th)
    for r in range(0, l - s + 1):
        k = r + s - m
        if k < 0 or k > l + s:
            continue
        tot = tot + comb(l - s, r) * comb(l + s, k) * (-1.0) ** (l - r - s) * \
            (np.cos(th / 2) / np.sin(th / 2)) ** (2 * r + s - m)
    return pre * np.sin(th / 2) ** (2 * l) * tot


@pytest.mark.parametrize("s", [1, 2, 3, -2])
def test_spin_lambda_matches_goldberg(s):
    """the spin recursion against the explicit sum, m >= 0, l <= 9"""
    th = np.linspace(0.05, np.pi - 0.05, 9)
    for m in range(0, 7):
        lam = S.lambda_spin(9, m, s, th)
        for l in range(10):
            assert np.allclose(lam[l], _goldberg(s, l, m, th), atol=1e-12), \
                (s, l, m)


def test_spin_zero_recursion_is_the_legendre_one():
    """s = 0, m >= 1 (at l = m = s = 0 the Wigner recursion divides by
    zero; only s >= 1 is used)"""
    th = np.linspace(0.05, np.pi - 0.05, 9)
    for m in range(1, 6):
        assert np.allclose(S.lambda_spin(12, m, 0, th)[m:],
                           S.lambda_m(12, m, th), atol=1e-13)


def test_spin_synthesis_is_the_two_sided_sum():
    """Q + i U = -sum over l and m = -l..l of (E + i B)_lm sY_lm, with the
    negative m from E_{l,-m} = (-1)^m conj(E_lm), by Goldberg's formula"""
    rng = np.random.default_rng(5)
    lmax, s = 7, 2
    ms = S.default_mstart(lmax)
    n = S.nalm(lmax, lmax)
    a = rng.uniform(-1, 1, (2, n)) + 1j * rng.uniform(-1, 1, (2, n))
    a[:, :lmax + 1].imag = 0
    loc = _loc(rng, 12)
    th, ph = loc[:, 0], loc[:, 1]
    P = np.zeros(len(th), complex)
    for l in range(lmax + 1):
        for M in range(-l, l + 1):
            e, b = a[0, ms[abs(M)] + l], a[1, ms[abs(M)] + l]
            if M < 0:
                e, b = (-1) ** M * np.conj(e), (-1) ** M * np.conj(b)
            P -= (e + 1j * b) * _goldberg(s, l, M, th) * np.exp(1j * M * ph) \
                if M >= 0 else \
                (e + 1j * b) * (-1.0) ** (s + M) * _goldberg(-s, l, -M, th) \
                * np.exp(1j * M * ph)
    q = S.synthesis_general_spin(a, lmax, loc, s, ms)
    assert np.allclose(q[0], P.real, atol=1e-12)
    assert np.allclose(q[1], P.imag, atol=1e-12)


def test_spin1_gradient_is_the_gradient():
    """spin 1, E_lm sqrt(l (l + 1)), B = 0 (the DERIV1 mode) gives
    (d f / d theta, d f / d phi / sin theta) of the spin-0 field f with the
    same a_lm: central differences"""
    rng = np.random.default_rng(0)
    lmax = 9
    ms = S.default_mstart(lmax)
    a = S.random_alm(lmax, lmax, rng)
    loc = rng.uniform(0.3, 2.8, (6, 2))
    loc[:, 1] *= 2
    h = 1e-6

    def f(x):
        return S.synthesis_general(a, lmax, x, ms)
    dth = (f(loc + [h, 0]) - f(loc - [h, 0])) / (2 * h)
    dph = (f(loc + [0, h]) - f(loc - [0, h])) / (2 * h) / np.sin(loc[:, 0])
    w = np.zeros(len(a))
    for m in range(lmax + 1):
        ls = np.arange(m, lmax + 1)
        w[ms[m] + ls] = np.sqrt(ls * (ls + 1.))
    q = S.synthesis_general_spin((a * w)[None, :], lmax, loc, 1, ms)
    assert np.allclose(q[0], dth, rtol=1e-7, atol=1e-7)
    assert np.allclose(q[1], dph, rtol=1e-7, atol=1e-7)


@pytest.mark.parametrize("s,ncomp", [(1, 2), (2, 2), (3, 1)])
def test_spin_adjoint_identity(s, ncomp):
    """<synthesis(a), r> = almdot over both components (m > 0 twice)"""
    rng = np.random.default_rng(s)
    lmax, mmax = 11, 8
    ms = S.default_mstart(lmax, mmax)
    n = S.nalm(lmax, mmax)
    a = rng.uniform(-1, 1, (ncomp, n)) + 1j * rng.uniform(-1, 1, (ncomp, n))
    a[:, :lmax + 1].imag = 0
    loc = _loc(rng, 30)
    r = rng.uniform(-1, 1, (2, 30))
    f = S.synthesis_general_spin(a, lmax, loc, s, ms)
    b = S.adjoint_synthesis_general_spin(r, lmax, loc, s, ncomp, ms, size=n)
    v1 = sum(S.almdot(a[c], b[c], lmax, ms) for c in range(ncomp))
    v2 = float(np.sum(r * f))
    assert abs(v1 - v2) <= 1e-12 * max(abs(v1), 1.0)


def test_a_planted_spin_sign_error_is_caught():
    """dropping the (-1)^s of the start value flips odd-spin functions,
    which the explicit sum detects"""
    th = np.linspace(0.2, 2.9, 5)
    lam = -S.lambda_spin(5, 2, 1, th)
    assert not np.allclose(lam[3], _goldberg(1, 3, 2, th), atol=1e-6)


@pytest.mark.parametrize("mode", ["STANDARD", "GRAD_ONLY", "DERIV1"])
def test_spin_matches_ducc0_itself(mode):
    """the spin restatement against ducc0's synthesis_general /
    adjoint_synthesis_general (oracle/_ref) at eps 1e-11"""
    from oracle import ref as R
    from oracle import core as O
    if not R.available():
        pytest.skip("reference library not built")
    rng = np.random.default_rng(2)
    lmax, mmax = 24, 20
    s = 1 if mode == "DERIV1" else 2
    ms = S.default_mstart(lmax, mmax)
    n = S.nalm(lmax, mmax)
    nc = 2 if mode == "STANDARD" else 1
    a = rng.uniform(-1, 1, (nc, n)) + 1j * rng.uniform(-1, 1, (nc, n))
    a[:, :lmax + 1].imag = 0
    w = np.ones(n)
    if mode == "DERIV1":
        for m in range(mmax + 1):
            ls = np.arange(m, lmax + 1)
            w[ms[m] + ls] = np.sqrt(ls * (ls + 1.))
    loc = _loc(rng, 300)
    f = R.sht_synthesis_general(a, lmax, loc, 1e-11, mstart=ms, spin=s,
                                mode=mode)
    q = S.synthesis_general_spin(a * w, lmax, loc, s, ms)
    assert O.l2error(f.ravel().astype(complex), q.ravel().astype(complex)) \
        <= 1e-10
    r = rng.uniform(-1, 1, (2, 300))
    b = R.sht_adjoint_synthesis_general(r, lmax, loc, 1e-11, n, mstart=ms,
                                        spin=s, mode=mode)
    bo = S.adjoint_synthesis_general_spin(r, lmax, loc, s, nc, ms, size=n) * w
    assert O.l2error(b.ravel(), bo.ravel()) <= 1e-10


# ------------------------------------------------- pseudo_analysis_general

def _noisy_problem(rng, lmax, spin, npt):
    n = S.nalm(lmax, lmax)
    nc = 1 if spin == 0 else 2
    a = rng.uniform(-1, 1, (nc, n)) + 1j * rng.uniform(-1, 1, (nc, n))
    a[:, :lmax + 1].imag = 0
    ofs = 0
    for s in range(spin):
        a[:, ofs:ofs + spin - s] = 0
        ofs += lmax + 1 - s
    loc = _loc(rng, npt)
    m = S.synthesis_general(a[0], lmax, loc)[None, :] if spin == 0 \
        else S.synthesis_general_spin(a, lmax, loc, spin)
    return a, loc, m + 1e-3 * rng.normal(size=m.shape)


@pytest.mark.parametrize("spin", [0, 2])
def test_pseudo_analysis_matches_ducc0_itself(spin):
    """scipy's LSMR on the direct sums against ducc0's
    pseudo_analysis_general (oracle/_ref): the same stopping reason,
    iteration count and residual, and the same a_lm"""
    from oracle import ref as R
    from oracle import core as O
    if not R.available():
        pytest.skip("reference library not built")
    rng = np.random.default_rng(1)
    lmax = 12
    # 2000 points for 169 coefficients: well conditioned, so that rounding
    # differences do not grow over the iterations (on 600 points a
    # perturbation of the map at 1e-15 moves the 25th iterate by 1e-5)
    a, loc, m = _noisy_problem(rng, lmax, spin, 2000)
    n = a.shape[1]
    for it in (5, 30):
        o = S.pseudo_analysis_general(m, lmax, loc, spin, it, 1e-10, size=n)
        r = R.sht_pseudo_analysis_general(m, lmax, loc, spin, it, 1e-10, n)
        assert o[1:3] == r[1:3]
        assert abs(o[3] - r[3]) <= 1e-6 * r[3]
        assert O.l2error(o[0].ravel(), r[0].ravel()) <= 1e-6


def test_pseudo_analysis_solves_the_normal_equations():
    """run to convergence on a full-rank problem, the solution satisfies
    A^T (A x - b) = 0 (numpy's dense least squares gives the same x)"""
    rng = np.random.default_rng(2)
    lmax = 6
    a, loc, m = _noisy_problem(rng, lmax, 0, 200)
    n = a.shape[1]
    x, istop, itn, rr, ar = S.pseudo_analysis_general(m, lmax, loc, 0, 200,
                                                      1e-13, size=n)
    assert istop in (1, 2)
    # the dense problem in the real coordinates of the oracle
    idx = S._layout(lmax, S.default_mstart(lmax), 1, 0)
    nx = len(S._alm_to_real(np.zeros((1, n), complex), idx))
    cols = []
    for k in range(nx):
        e = np.zeros(nx)
        e[k] = 1
        cols.append(S.synthesis_general(S._real_to_alm(e, idx, 1, n)[0], lmax,
                                        loc))
    Am = np.stack(cols, 1)
    xr = np.linalg.lstsq(Am, m.ravel(), rcond=None)[0]
    assert np.allclose(S._alm_to_real(x, idx), xr, atol=1e-9)
