"""Pinning the CPU oracle (oracle/oracle.c) before trusting it.

Each restated function is checked against what the mathematics fixes
(closed forms, textbook library routines, brute force) and against the
golden fixtures written from the compiled reference ducc0
(tests/golden/make_golden.py).  Where oracle/_ref was built in this
container, further random cases are compared live."""
import json
import os

import numpy as np
import pytest

from oracle import core as O
from oracle import ref as R

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
have_ref = R.available()


# ------------------------------------------------------------ closed forms
def test_good_size_brute_force():
    def smooth(x):
        for p in (2, 3, 5, 7, 11):
            while x % p == 0:
                x //= p
        return x == 1
    for n in list(range(1, 400)) + [1693, 2049, 3387, 10007, 65537]:
        g = O.good_size(n)
        if n <= 12:
            assert g == n
            continue
        m = n
        while not smooth(m):
            m += 1
        assert g == m, n


def test_gl_nodes_vs_numpy():
    for n in (2, 5, 8, 14, 26, 52):
        x, w = O.gl_nodes(n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        m = (n + 1) // 2
        # ducc0 order: non-negative nodes from the smallest
        sel = np.argsort(xr)[n - m:]
        assert np.allclose(x, xr[sel], atol=1e-15, rtol=0)
        assert np.allclose(w, wr[sel], atol=1e-14, rtol=1e-13)


def test_es_kernel_closed_form():
    for beta, e0 in ((10.0, 0.5), (23.4, 0.52)):
        for v in (-1.0, -0.3, 0.0, 0.7, 1.0, 1.2):
            t = (1 - v) * (1 + v)
            ref = np.exp(beta * (t ** e0 - 1)) if t >= 0 else 0.0
            assert O.esk(v, beta, e0) == pytest.approx(ref, rel=1e-15, abs=0)


@pytest.mark.parametrize("row", [0, 100, 400, 700, 1000, 1300])
def test_kernel_coeffs_vs_chebyshev(row):
    """Each piece equals numpy's Chebyshev interpolant of psi at the D+1
    Chebyshev nodes, converted to monomials of the local coordinate."""
    r = O.db_row(row)
    W = r["W"]
    D = W + 3
    c = O.kernel_coeffs(W, r["beta"], r["e0"])
    beta = r["beta"] * W
    for p in range(W):
        lo, hi = -1 + 2 * p / W, -1 + 2 * (p + 1) / W
        f = lambda t: np.array([O.esk(0.5 * (hi - lo) * ti + 0.5 * (hi + lo),
                                       beta, r["e0"]) for ti in t])
        cheb = np.polynomial.chebyshev.Chebyshev.interpolate(f, D)
        mono = np.polynomial.chebyshev.cheb2poly(cheb.coef)
        mine = c[::-1, p]
        xs = np.linspace(-1, 1, 41)
        assert np.allclose(np.polyval(mine[::-1], xs),
                           np.polynomial.polynomial.polyval(xs, mono),
                           atol=1e-13, rtol=0)
    # the pieces approximate psi itself to about the tabulated accuracy
    xs = np.linspace(-1, 1, 2001)
    vals = []
    for xx in xs:
        xrel = W * 0.5 * (xx + 1)
        nth = min(int(xrel), W - 1)
        loc = ((xrel - nth) - 0.5) * 2
        vals.append(np.polyval(c[:, nth], loc))
    exact = np.array([O.esk(x, beta, r["e0"]) for x in xs])
    assert np.max(np.abs(np.array(vals) - exact)) < 10 * r["epsilon"] + 1e-12


def test_corfac_vs_direct_quadrature():
    for row in (50, 600, 1250):
        r = O.db_row(row)
        W = r["W"]
        beta = r["beta"] * W
        p = int(1.5 * W) + 2
        xg, wg = np.polynomial.legendre.leggauss(2 * p)
        pos = xg >= 0
        x, w = xg[pos], 2 * wg[pos]
        psi = np.array([O.esk(t, beta, r["e0"]) for t in x])
        n, dx = 100, 1 / 317.
        k = np.arange(n)[:, None]
        ref = 1 / np.sum(w * psi * W * 0.5 * np.cos(np.pi * W * x * k * dx),
                         axis=1)
        got = O.corfac(W, r["beta"], r["e0"], n, dx)
        assert np.allclose(got, ref, rtol=1e-13, atol=0)


def test_fft_vs_numpy():
    rng = np.random.default_rng(0)
    for shape in ((16,), (462,), (30, 22), (12, 14, 10), (77,)):
        a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        assert np.allclose(O.fft_c2c(a, True), np.fft.fftn(a),
                           atol=1e-11 * np.sqrt(a.size))
        assert np.allclose(O.fft_c2c(a, False), np.fft.ifftn(a) * a.size,
                           atol=1e-11 * np.sqrt(a.size))


def test_l2error_closed_form():
    a = np.array([1 + 1j, 2, 0])
    b = np.array([1 + 1j, 2 + 1j, 0])
    assert O.l2error(a, b) == pytest.approx(np.sqrt(1 / 7), rel=1e-15)
    assert O.l2error(a, a) == 0.0


@pytest.mark.parametrize("fwd", [True, False])
def test_nudft_single_point_closed_form(fwd):
    """M=1: nu2u output is c*exp(isign i k x) exactly."""
    x = np.array([[0.7, -2.1]])
    c = np.array([0.3 - 0.4j])
    shape = (6, 5)
    out = O.nudft_type1(x, c, shape, fwd)
    k0 = np.arange(6) - 3
    k1 = np.arange(5) - 2
    s = -1 if fwd else 1
    ref = c[0] * np.exp(s * 1j * (k0[:, None] * x[0, 0] + k1[None, :] * x[0, 1]))
    assert np.allclose(out, ref, atol=1e-15)


def test_nudft_shift_theorem():
    """Translating coords by delta multiplies type-1 output by
    exp(isign i k delta)."""
    rng = np.random.default_rng(3)
    x = rng.uniform(-3, 3, (20, 1))
    c = rng.standard_normal(20) + 1j * rng.standard_normal(20)
    a = O.nudft_type1(x, c, (9,), True)
    b = O.nudft_type1(x + 0.25, c, (9,), True)
    k = np.arange(9) - 4
    assert np.allclose(b, a * np.exp(-1j * k * 0.25), atol=1e-13)


@pytest.mark.parametrize("shape", [(37,), (20, 31), (8, 10, 12)])
@pytest.mark.parametrize("eps", [1e-2, 1e-5, 1e-10])
def test_oracle_nufft_within_eps(shape, eps):
    rng = np.random.default_rng(5)
    M = 300
    nd = len(shape)
    c = rng.uniform(-np.pi, np.pi, (M, nd))
    p = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    g = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    mul = {1: 4, 2: 2, 3: 1}[nd]
    out = O.nu2u(c, p, shape, True, eps)
    assert O.l2error(out, O.nudft_type1(c, p, shape, True)) <= mul * eps
    m = O.u2nu(c, g, False, eps)
    assert O.l2error(m, O.nudft_type2(c, g, False)) <= mul * eps
    # adjointness of the approximations themselves
    lhs = np.vdot(p, O.u2nu(c, g, False, eps))
    rhs = np.vdot(out, g)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1)


def test_planted_sign_error_is_caught():
    """A plausible mistake (wrong exponent sign) fails the NUDFT check."""
    rng = np.random.default_rng(9)
    c = rng.uniform(-np.pi, np.pi, (50, 2))
    p = rng.standard_normal(50) + 1j * rng.standard_normal(50)
    good = O.nu2u(c, p, (16, 16), True, 1e-8)
    wrong = O.nu2u(c, p, (16, 16), False, 1e-8)
    ref = O.nudft_type1(c, p, (16, 16), True)
    assert O.l2error(good, ref) < 1e-7
    assert O.l2error(wrong, ref) > 0.1


# ------------------------------------------------------------ golden fixtures
def _gold(name):
    return os.path.join(GOLD, name)


def test_params_vs_ducc0_golden():
    g = json.load(open(_gold("params.json")))
    for nd in (1, 2, 3):
        for s in (True, False):
            assert O.best_epsilon(nd, s) == g["best_epsilon"][
                f"{nd}{'s' if s else 'd'}"]
    n = 0
    for r in g["requests"]:
        vlen = g["vlen_single"] if r["single"] else g["vlen_double"]
        k, nov = O.find_params(r["eps"], r["shape"], r["M"], r["gridding"],
                               r["single"], vlen=vlen, nthreads=r["nthreads"],
                               sigma_min=r["sigma_min"],
                               sigma_max=r["sigma_max"])
        assert k == r["kidx"], r
        if k >= 0:
            assert list(nov) == r["nover"], r
            n += 1
    assert n > 300


def test_kernels_vs_ducc0_golden():
    z = np.load(_gold("kernels.npz"))
    for k in z["rows"]:
        r = O.db_row(int(k))
        c = O.kernel_coeffs(r["W"], r["beta"], r["e0"])
        ref = z[f"coeff_{k}"]
        # high-order monomial coefficients carry cancellation; compare the
        # polynomials on the piece instead
        xs = np.linspace(-1, 1, 33)
        for p in range(r["W"]):
            assert np.allclose(np.polyval(c[:, p], xs),
                               np.polyval(ref[:, p], xs), atol=1e-14, rtol=0)
        # k*dx <= 0.5 is the range a transform uses (k <= n/2, dx = 1/nover);
        # there the sum has lost at most ~1e4 of relative conditioning
        cf = O.corfac(r["W"], r["beta"], r["e0"], 300, 1 / 517.)
        assert np.allclose(cf[:259], z[f"corfac_{k}"][:259], rtol=1e-11,
                           atol=0)


def test_sort_vs_ducc0_golden():
    """oracle keys (with ducc0's own tile sizes and 3-D refinement) sorted
    stably reproduce ducc0's coord_idx exactly."""
    z = np.load(_gold("sort.npz"))
    i = 0
    while f"coord_{i}" in z:
        c = z[f"coord_{i}"]
        nd, single, supp = [int(v) for v in z[f"meta_{i}"][:3]]
        nover = tuple(int(v) for v in z[f"meta_{i}"][3:])
        l2 = {1: 9, 2: 5 if single else 4, 3: 4}[nd]
        keys = O.tile_keys(c, supp, nover, 2 * np.pi, l2,
                           lsq2=-1 if nd == 3 else None)
        idx = np.argsort(keys, kind="stable")
        assert np.array_equal(idx, z[f"idx_{i}"]), i
        assert np.array_equal(O.sort_stable(keys), z[f"idx_{i}"])
        i += 1
    assert i == 5


def test_nufft_vs_ducc0_golden():
    z = np.load(_gold("nufft.npz"))
    cases = json.loads(str(z["cases"]))
    gp = json.load(open(_gold("params.json")))
    for i, cs in enumerate(cases):
        c = z[f"coord_{i}"]
        p = z[f"points_{i}"]
        g = z[f"grid_{i}"]
        vlen = gp["vlen_single"] if cs["single"] else gp["vlen_double"]
        kw = dict(single=cs["single"], vlen=vlen, nthreads=1,
                  period=cs["periodicity"], fft_order=cs["fft_order"])
        a = O.nu2u(c, p, tuple(cs["shape"]), cs["forward"], cs["eps"], **kw)
        b = O.u2nu(c, g, not cs["forward"], cs["eps"], **kw)
        # the oracle computes in FP64; ducc0 in the value precision
        tol = 1e-12 if not cs["single"] else 3e-6
        assert O.l2error(a, z[f"nu2u_{i}"]) < tol, (i, cs)
        assert O.l2error(b, z[f"u2nu_{i}"]) < tol, (i, cs)


# --------------------------------------------------- live reference (if built)
@pytest.mark.skipif(not have_ref, reason="oracle/_ref not built here")
def test_live_params_vs_ducc0():
    rng = np.random.default_rng(11)
    for _ in range(300):
        nd = int(rng.integers(1, 4))
        shape = tuple(int(x) for x in rng.integers(1, 500, nd))
        single = bool(rng.integers(0, 2))
        eps = 10 ** rng.uniform(-14 if not single else -6.5, -1)
        M = int(10 ** rng.uniform(0, 8))
        nth = int(rng.integers(1, 9))
        try:
            r = R.find_params(eps, shape, M, True, single, nth)
        except RuntimeError:
            r = (-1, ())
        o = O.find_params(eps, shape, M, True, single, vlen=R.vlen(single),
                          nthreads=nth)
        assert r[0] == o[0] and (r[0] < 0 or r[1] == o[1])


@pytest.mark.skipif(not have_ref, reason="oracle/_ref not built here")
@pytest.mark.parametrize("nd,shape", [(1, (300,)), (2, (90, 70)),
                                      (3, (20, 24, 18))])
def test_live_nufft_vs_ducc0(nd, shape):
    rng = np.random.default_rng(21)
    M = 2000
    c = rng.uniform(-np.pi, np.pi, (M, nd))
    p = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    g = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    for eps in (1e-4, 1e-9, 1e-13):
        a = O.nu2u(c, p, shape, False, eps, vlen=R.vlen(False), nthreads=1)
        b = R.nu2u(p, c, False, eps, shape)
        assert O.l2error(a, b) < 1e-12
        a2 = O.u2nu(c, g, True, eps, vlen=R.vlen(False), nthreads=1)
        b2 = R.u2nu(g, c, True, eps)
        assert O.l2error(a2, b2) < 1e-12
