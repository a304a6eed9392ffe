"""CPU restatement of ducc0's spin-0 general spherical harmonic synthesis
and its adjoint (test infrastructure only: used by tests/ as the checker,
never by the library).

synthesis_general (src/ducc0/sht/sht.cc:2926-2960) evaluates the real field
    f(theta, phi) = sum_{l, m} a_lm Y_lm(theta, phi)
                  = sum_{m >= 0} w_m Re(e^{i m phi} sum_l a_lm lambda_lm(theta)),
w_0 = 1, w_m = 2 (a_{l,-m} = (-1)^m conj(a_lm)), at arbitrary points, with
the orthonormal associated Legendre functions lambda_lm including the
Condon-Shortley phase (sht.cc:134-136, 552: lambda_mm = (-1)^m mfac_m
sin^m theta, mfac_m = prod_{i<=m} sqrt((2i+1)/(2i)) / sqrt(4 pi)) and the
three-term recursion of sht.cc:204-218 (eps_l = sqrt((l^2-m^2)/(4l^2-1)),
lambda_{l+1} = (cos theta lambda_l - eps_l lambda_{l-1}) / eps_{l+1}).
Here the sums are taken directly (no NUFFT): exact up to rounding, for the
small sizes of the tests.  a_lm layout: index mstart[m] + l * lstride
(python/sht_pymod.cc:124-140).

The adjoint (sht.cc:2962-2993) with respect to the map's real inner product
and ducc0's a_lm inner product (m = 0 once, m > 0 twice, the compress_alm
of python/test/test_sht.py:43-52) is
    b_lm = sum_i r_i lambda_lm(theta_i) e^{-i m phi_i}   (b_l0 real).
Pinned in tests/test_sht_oracle.py against scipy.special.sph_harm_y, the
orthonormality of Y_lm under Gauss-Legendre quadrature, and the adjoint
identity."""
import numpy as np


def default_mstart(lmax, mmax=None):
    """sht_pymod.cc:124-134"""
    mmax = lmax if mmax is None else mmax
    ms = np.zeros(mmax + 1, np.int64)
    idx = 0
    for m in range(mmax + 1):
        ms[m] = idx
        idx += lmax - m
    return ms


def nalm(lmax, mmax):
    return ((mmax + 1) * (mmax + 2)) // 2 + (mmax + 1) * (lmax - mmax)


def lambda_m(lmax, m, theta):
    """lambda_lm(cos theta) for l = m..lmax (rows) at the given theta"""
    x = np.cos(theta)
    st = np.sin(theta)
    mf = 1.0 / np.sqrt(4 * np.pi)
    for i in range(1, m + 1):
        mf *= np.sqrt((2 * i + 1.0) / (2 * i))
    out = np.zeros((lmax + 1 - m, len(theta)))
    with np.errstate(under="ignore"):
        lam = ((-1) ** m) * mf * st ** m
    lamp = np.zeros_like(lam)

    def eps(l):
        return 0.0 if l <= m else np.sqrt((l * l - m * m) / (4.0 * l * l - 1))
    for l in range(m, lmax + 1):
        out[l - m] = lam
        e1 = eps(l + 1)
        nl = (x * lam - eps(l) * lamp) / e1
        lamp, lam = lam, nl
    return out


def synthesis_general(alm, lmax, loc, mstart=None, lstride=1):
    """map (npoints,) of the real field with coefficients alm (m >= 0)"""
    alm = np.asarray(alm).ravel()
    ms = default_mstart(lmax) if mstart is None else np.asarray(mstart)
    th, ph = loc[:, 0], loc[:, 1]
    res = np.zeros(len(th))
    for m in range(len(ms)):
        lam = lambda_m(lmax, m, th)
        ls = np.arange(m, lmax + 1)
        a = alm[ms[m] + ls * lstride]
        leg = a @ lam  # sum over l
        w = 1.0 if m == 0 else 2.0
        res += w * np.real(leg * np.exp(1j * m * ph))
    return res


def adjoint_synthesis_general(mapv, lmax, loc, mstart=None, lstride=1,
                              size=None):
    """a_lm (complex) from the real values mapv at loc"""
    mapv = np.asarray(mapv, np.float64).ravel()
    ms = default_mstart(lmax) if mstart is None else np.asarray(mstart)
    th, ph = loc[:, 0], loc[:, 1]
    n = size
    if n is None:
        n = max(int(ms[m]) + lmax * lstride for m in range(len(ms))) + 1
    out = np.zeros(n, np.complex128)
    for m in range(len(ms)):
        lam = lambda_m(lmax, m, th)
        v = lam @ (mapv * np.exp(-1j * m * ph))
        if m == 0:
            v = v.real.astype(np.complex128)
        ls = np.arange(m, lmax + 1)
        out[ms[m] + ls * lstride] = v
    return out


def almdot(a, b, lmax, mstart=None, lstride=1):
    """ducc0's a_lm inner product for real fields: m = 0 entries once
    (real parts), m > 0 entries twice (test_sht.py:43-52)"""
    ms = default_mstart(lmax) if mstart is None else np.asarray(mstart)
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    s = 0.0
    for m in range(len(ms)):
        ls = np.arange(m, lmax + 1)
        ia = ms[m] + ls * lstride
        t = np.real(np.conj(a[ia]) * b[ia])
        s += np.sum(t) if m == 0 else 2 * np.sum(t)
    return s


def random_alm(lmax, mmax, rng):
    """test_sht.py:31-40: uniform real / imaginary parts in [-1, 1), m = 0
    real (default layout)"""
    n = nalm(lmax, mmax)
    res = rng.uniform(-1., 1., n) + 1j * rng.uniform(-1., 1., n)
    res[0:lmax + 1].imag = 0.
    return res


# ---------------------------------------------------------------- spin > 0
# Spin-s fields (sht.cc:803-1010 computes them by its own recursion): with
# the spin-weighted harmonics sY_lm = lambda^s_lm(theta) e^{i m phi}
# (Goldberg's convention),
#   Q + i U = - sum_{l >= s, m = -l..l} (E_lm + i B_lm) sY_lm,
#   E_{l,-m} = (-1)^m conj(E_lm), B likewise,
# which is what ducc0 computes for spin s in STANDARD mode (GRAD_ONLY: B = 0;
# checked against ducc0 itself in tests/test_sht_oracle.py).
# lambda^s_lm for m >= 0 by the three-term recursion in l of the Wigner
# functions d^l_{m, m'} with m' = -s, lambda = (-1)^s sqrt((2l+1)/(4 pi))
# d^l_{m,-s}, started at l0 = max(m, |s|) from the closed forms of
# d^{l0}_{l0, m'} and its symmetries.

def _log_binom_sqrt(j, a):
    """log sqrt((2j)! / ((j+a)! (j-a)!))"""
    from math import lgamma
    return 0.5 * (lgamma(2 * j + 1) - lgamma(j + a + 1) - lgamma(j - a + 1))


def lambda_spin(lmax, m, s, theta):
    """lambda^s_lm(theta), rows l = 0..lmax (zero below max(m, |s|))"""
    th = np.asarray(theta, np.float64)
    mp = -s
    x = np.cos(th)
    c, sn = np.cos(th / 2), np.sin(th / 2)
    l0 = max(m, abs(mp))
    out = np.zeros((lmax + 1, len(th)))
    if l0 > lmax:
        return out
    with np.errstate(under="ignore", divide="ignore"):
        lc, ls = np.log(c), np.log(sn)
        if m >= abs(mp):
            sgn = (-1.0) ** (m - mp)
            lg = _log_binom_sqrt(m, mp) + (m + mp) * lc + (m - mp) * ls
        elif mp > 0:
            sgn = 1.0
            lg = _log_binom_sqrt(l0, m) + (l0 + m) * lc + (l0 - m) * ls
        else:
            sgn = (-1.0) ** (l0 + m)
            lg = _log_binom_sqrt(l0, m) + (l0 - m) * lc + (l0 + m) * ls
        d0 = sgn * np.exp(lg)
    N = lambda l: np.sqrt((2 * l + 1) / (4 * np.pi))
    R = lambda l: np.sqrt(max(l * l - m * m, 0) * max(l * l - mp * mp, 0))
    lam = (-1.0) ** s * N(l0) * d0
    lamp = np.zeros_like(lam)
    for l in range(l0, lmax + 1):
        out[l] = lam
        if l == lmax:
            break
        r1 = R(l + 1)
        a = N(l + 1) / N(l) * (2 * l + 1) * (l + 1) / r1
        b = -N(l + 1) / N(l) * (2 * l + 1) * m * mp / (l * r1)
        cc = N(l + 1) / N(l - 1) * (l + 1) * R(l) / (l * r1)
        lamp, lam = lam, (a * x + b) * lam - cc * lamp
    return out


def synthesis_general_spin(alm, lmax, loc, spin, mstart=None, lstride=1):
    """(Q, U) (2, npoints) from alm (2, nalm) = (E, B), or (1, nalm) = E
    alone (GRAD_ONLY)"""
    alm = np.atleast_2d(np.asarray(alm))
    ms = default_mstart(lmax) if mstart is None else np.asarray(mstart)
    th, ph = loc[:, 0], loc[:, 1]
    P = np.zeros(len(th), np.complex128)
    for m in range(len(ms)):
        ls = np.arange(lmax + 1)
        E = alm[0, ms[m] + ls * lstride]
        B = alm[1, ms[m] + ls * lstride] if alm.shape[0] > 1 else 0 * E
        lp = lambda_spin(lmax, m, spin, th)
        a = E + 1j * B
        P -= (a @ lp) * np.exp(1j * m * ph)
        if m > 0:
            lmn = (-1.0) ** (spin + m) * lambda_spin(lmax, m, -spin, th)
            an = (-1.0) ** m * (np.conj(E) + 1j * np.conj(B))
            P -= (an @ lmn) * np.exp(-1j * m * ph)
    return np.stack([P.real, P.imag])


def adjoint_synthesis_general_spin(mapv, lmax, loc, spin, ncomp=2,
                                   mstart=None, lstride=1, size=None):
    """the adjoint of synthesis_general_spin with respect to the real inner
    products of (Q, U) and of the a_lm (m > 0 twice): (ncomp, nalm)"""
    mapv = np.asarray(mapv, np.float64)
    r = mapv[0] + 1j * mapv[1]
    ms = default_mstart(lmax) if mstart is None else np.asarray(mstart)
    th, ph = loc[:, 0], loc[:, 1]
    n = size
    if n is None:
        n = max(int(ms[m]) + lmax * lstride for m in range(len(ms))) + 1
    out = np.zeros((ncomp, n), np.complex128)
    for m in range(len(ms)):
        ls = np.arange(m, lmax + 1)
        lp = lambda_spin(lmax, m, spin, th)[m:]
        lm = lambda_spin(lmax, m, -spin, th)[m:]
        hp = lp @ (r * np.exp(-1j * m * ph))
        hm = lm @ (r * np.exp(1j * m * ph))
        sg = (-1.0) ** spin
        bE = (-hp - sg * np.conj(hm)) / 2
        bB = 1j * (hp - sg * np.conj(hm)) / 2
        out[0, ms[m] + ls * lstride] = bE
        if ncomp > 1:
            out[1, ms[m] + ls * lstride] = bB
    return out


# ------------------------------------------------- pseudo_analysis_general
# ducc0 (sht.cc:3003-3058) solves min |synthesis(a) - map| by LSMR with the
# adjoint above, the a_lm norm counting m > 0 twice.  Here the a_lm are
# mapped to real coordinates (m = 0: the real part; m > 0: sqrt(2) times
# the real and imaginary parts), in which that norm is the Euclidean one
# and the adjoint is the transpose, and scipy's LSMR (the same published
# algorithm, Fong & Saunders 2011) runs on the direct sums.

def _layout(lmax, ms, lstride, spin):
    idx = []
    for m in range(len(ms)):
        for l in range(m, lmax + 1):
            idx.append((m, int(ms[m]) + l * lstride))
    return idx


def _alm_to_real(a, idx):
    out = []
    for c in range(a.shape[0]):
        for m, ix in idx:
            if m == 0:
                out.append(a[c, ix].real)
            else:
                out.append(np.sqrt(2.) * a[c, ix].real)
                out.append(np.sqrt(2.) * a[c, ix].imag)
    return np.asarray(out)


def _real_to_alm(x, idx, ncomp, size):
    a = np.zeros((ncomp, size), np.complex128)
    k = 0
    for c in range(ncomp):
        for m, ix in idx:
            if m == 0:
                a[c, ix] = x[k]
                k += 1
            else:
                a[c, ix] = (x[k] + 1j * x[k + 1]) / np.sqrt(2.)
                k += 2
    return a


def pseudo_analysis_general(mapv, lmax, loc, spin, maxiter, epsilon,
                            mstart=None, lstride=1, size=None):
    """(alm (ncomp, size), istop, itn, |r| / |b|, |A^T r| / (|A| |r|))"""
    from scipy.sparse.linalg import LinearOperator, lsmr
    mapv = np.atleast_2d(np.asarray(mapv, np.float64))
    ncomp = 1 if spin == 0 else 2
    ms = default_mstart(lmax) if mstart is None else np.asarray(mstart)
    n = size
    if n is None:
        n = max(int(ms[m]) + lmax * lstride for m in range(len(ms))) + 1
    idx = _layout(lmax, ms, lstride, spin)
    npt = loc.shape[0]

    def mv(x):
        a = _real_to_alm(np.asarray(x).ravel(), idx, ncomp, n)
        if spin == 0:
            return synthesis_general(a[0], lmax, loc, ms, lstride)
        return synthesis_general_spin(a, lmax, loc, spin, ms,
                                      lstride).ravel()

    def rmv(y):
        y = np.asarray(y).ravel().reshape(ncomp, npt)
        if spin == 0:
            a = adjoint_synthesis_general(y[0], lmax, loc, ms, lstride,
                                          size=n)[None, :]
        else:
            a = adjoint_synthesis_general_spin(y, lmax, loc, spin, ncomp,
                                               ms, lstride, size=n)
        return _alm_to_real(a, idx)

    nx = len(_alm_to_real(np.zeros((ncomp, n), complex), idx))
    A = LinearOperator((ncomp * npt, nx), matvec=mv, rmatvec=rmv,
                       dtype=np.float64)
    b = mapv.ravel()
    x, istop, itn, normr, normar, normA, condA, normx = lsmr(
        A, b, damp=0.0, atol=1e-14 * np.sqrt(npt), btol=epsilon,
        conlim=1e8, maxiter=maxiter)
    nb = np.linalg.norm(b)
    return (_real_to_alm(x, idx, ncomp, n), istop, itn, normr / nb,
            normar / (normA * normr) if normA * normr else 0.0)
