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
