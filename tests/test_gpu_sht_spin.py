"""ducc0.sht.synthesis_general / adjoint_synthesis_general for spin > 0 on
the GPU (modes STANDARD, GRAD_ONLY, DERIV1) against the oracle's direct
sums (oracle/sht.py: spin-weighted harmonics pinned to Goldberg's explicit
formula, to the gradient of the spin-0 field and to ducc0 itself), the
reference's adjointness test with spin (python/test/test_sht.py:234-262,
rtol 1e-9), its accuracy bound (l2error within epsilon) and ducc0's own
synthesis at lmax 400."""
import numpy as np
import pytest
import torch

from oracle import core as O
from oracle import sht as S

pytestmark = pytest.mark.gpu


def _loc(rng, n):
    loc = rng.uniform(0., 1., (n, 2))
    loc[:, 0] *= np.pi
    loc[:, 1] *= 2 * np.pi
    return loc


def _random_alm(lmax, mmax, spin, ncomp, rng):
    """test_sht.py:31-40: uniform parts in [-1, 1), m = 0 real, l < spin
    zero"""
    n = S.nalm(lmax, mmax)
    res = rng.uniform(-1., 1., (ncomp, n)) + \
        1j * rng.uniform(-1., 1., (ncomp, n))
    res[:, 0:lmax + 1].imag = 0.
    ofs = 0
    for s in range(spin):
        res[:, ofs:ofs + spin - s] = 0.
        ofs += lmax + 1 - s
    return res


def _deriv1_weight(lmax, mmax):
    ms = S.default_mstart(lmax, mmax)
    w = np.ones(S.nalm(lmax, mmax))
    for m in range(mmax + 1):
        ls = np.arange(m, lmax + 1)
        w[ms[m] + ls] = np.sqrt(ls * (ls + 1.))
    return w


CASES = [
    # lmax, mmax, spin, npoints, eps, complex dtype, mode
    (2, 2, 2, 50, 1e-10, np.complex128, "STANDARD"),
    (5, 5, 1, 300, 1e-10, np.complex128, "STANDARD"),
    (10, 8, 2, 1000, 1e-7, np.complex128, "STANDARD"),
    (16, 16, 3, 800, 1e-11, np.complex128, "GRAD_ONLY"),
    (40, 30, 2, 1500, 1e-5, np.complex128, "STANDARD"),
    (32, 32, 1, 1000, 1e-9, np.complex128, "DERIV1"),
    (64, 40, 2, 2000, 1e-4, np.complex64, "STANDARD"),
    (60, 60, 5, 1000, 1e-8, np.complex128, "STANDARD"),
    (50, 50, 2, 700, 1e-5, np.complex64, "GRAD_ONLY"),
    (3, 3, 5, 100, 1e-10, np.complex128, "STANDARD"),  # spin > lmax: zero
]


def _ids(c):
    return f"l{c[0]}m{c[1]}s{c[2]}n{c[3]}e{c[4]:g}{np.dtype(c[5]).name}{c[6]}"


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_spin_synthesis_vs_direct_sum(gpu, case):
    from ducc_amd import sht
    lmax, mmax, spin, npt, eps, cdt, mode = case
    print("spin synthesis", _ids(case), flush=True)
    rng = np.random.default_rng(lmax + 3 * spin + npt)
    nc = 2 if mode == "STANDARD" else 1
    a = _random_alm(lmax, mmax, spin, nc, rng).astype(cdt)
    loc = _loc(rng, npt)
    got = sht.synthesis_general(alm=a, spin=spin, lmax=lmax, mmax=mmax,
                                loc=loc, epsilon=eps, mode=mode)
    assert got.shape == (2, npt)
    assert got.dtype == (np.float64 if cdt == np.complex128 else np.float32)
    w = _deriv1_weight(lmax, mmax) if mode == "DERIV1" else 1.0
    ref = S.synthesis_general_spin(a.astype(np.complex128) * w, lmax, loc,
                                   spin, S.default_mstart(lmax, mmax))
    if spin > lmax:
        assert np.all(got == 0)
        return
    assert O.l2error(got.ravel().astype(np.complex128),
                     ref.ravel().astype(np.complex128)) <= eps


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_spin_adjoint_vs_direct_sum(gpu, case):
    from ducc_amd import sht
    lmax, mmax, spin, npt, eps, cdt, mode = case
    print("spin adjoint", _ids(case), flush=True)
    rng = np.random.default_rng(11 + lmax + spin + npt)
    loc = _loc(rng, npt)
    rdt = np.float64 if cdt == np.complex128 else np.float32
    r = rng.uniform(-0.5, 0.5, (2, npt)).astype(rdt)
    nc = 2 if mode == "STANDARD" else 1
    got = sht.adjoint_synthesis_general(map=r, spin=spin, lmax=lmax,
                                        mmax=mmax, loc=loc, epsilon=eps,
                                        mode=mode)
    assert got.dtype == cdt and got.shape == (nc, S.nalm(lmax, mmax))
    w = _deriv1_weight(lmax, mmax) if mode == "DERIV1" else 1.0
    ref = S.adjoint_synthesis_general_spin(
        r, lmax, loc, spin, nc, S.default_mstart(lmax, mmax),
        size=S.nalm(lmax, mmax)) * w
    if spin > lmax:
        assert np.all(got == 0)
        return
    assert O.l2error(got.ravel().astype(np.complex128), ref.ravel()) <= eps
    # the l < spin entries and the imaginary parts at m = 0 are exactly zero
    ms = S.default_mstart(lmax, mmax)
    for m in range(min(spin, mmax + 1)):
        assert np.all(got[:, ms[m] + np.arange(m, spin)] == 0)
    assert np.all(got[:, :lmax + 1].imag == 0)


@pytest.mark.parametrize("spin", [1, 2])
@pytest.mark.parametrize("npix", [33, 200])
@pytest.mark.parametrize("lmmax", [(10, 8), (100, 100)])
def test_adjointness_general_spin(gpu, lmmax, npix, spin):
    """test_sht.py:234-262 for spin 1, 2: STANDARD, then GRAD_ONLY"""
    from ducc_amd import sht
    rng = np.random.default_rng(48)
    lmax, mmax = lmmax
    eps = 1e-5
    slm1 = _random_alm(lmax, mmax, spin, 2, rng)
    loc = _loc(rng, npix)
    points2 = rng.uniform(-0.5, 0.5, (loc.shape[0], 2)).T.copy()
    ms = S.default_mstart(lmax, mmax)
    for mode, nc in (("STANDARD", 2), ("GRAD_ONLY", 1)):
        points1 = sht.synthesis_general(lmax=lmax, mmax=mmax, alm=slm1[:nc],
                                        loc=loc, spin=spin, epsilon=eps,
                                        mode=mode)
        slm2 = sht.adjoint_synthesis_general(lmax=lmax, mmax=mmax,
                                             map=points2, loc=loc, spin=spin,
                                             epsilon=eps, mode=mode)
        v1 = sum(S.almdot(slm1[c], slm2[c], lmax, ms) for c in range(nc))
        v2 = float(np.vdot(points2, points1))
        np.testing.assert_allclose(v1, v2, rtol=1e-9)


def test_spin_layout_stride_and_torch(gpu):
    """custom mstart / lstride, a component-strided map and torch device
    arrays give the numpy results"""
    from ducc_amd import sht
    rng = np.random.default_rng(19)
    lmax, mmax, spin, eps, lstride = 30, 20, 2, 1e-10, 2
    ms = S.default_mstart(lmax, mmax) * lstride + 3
    n = int(max(ms[m] + lmax * lstride for m in range(mmax + 1))) + 1
    a = np.zeros((2, n), np.complex128)
    for m in range(mmax + 1):
        for l in range(max(m, spin), lmax + 1):
            a[:, ms[m] + l * lstride] = rng.uniform(-1, 1, 2) + \
                (1j * rng.uniform(-1, 1, 2) if m else 0)
    loc = _loc(rng, 700)
    ref = S.synthesis_general_spin(a, lmax, loc, spin, ms, lstride)
    got = sht.synthesis_general(alm=a, spin=spin, lmax=lmax, mstart=ms,
                                lstride=lstride, loc=loc, epsilon=eps)
    assert O.l2error(got.ravel().astype(complex),
                     ref.ravel().astype(complex)) <= eps
    mp = torch.zeros((2, 2 * len(loc)), dtype=torch.float64, device="cuda")
    out = mp[:, ::2]
    res = sht.synthesis_general(alm=torch.from_numpy(a).cuda(), spin=spin,
                                lmax=lmax, mstart=ms, lstride=lstride,
                                loc=torch.from_numpy(loc).cuda(),
                                epsilon=eps, map=out)
    assert res is out
    np.testing.assert_allclose(out.cpu().numpy(), got, rtol=0, atol=1e-13)
    assert torch.all(mp[:, 1::2] == 0)
    r = rng.uniform(-1, 1, (2, len(loc)))
    at = torch.full((2, n), 7.0 + 0j, dtype=torch.complex128, device="cuda")
    sht.adjoint_synthesis_general(map=torch.from_numpy(r).cuda(), spin=spin,
                                  lmax=lmax, mstart=ms, lstride=lstride,
                                  loc=torch.from_numpy(loc).cuda(),
                                  epsilon=eps, alm=at)
    bref = S.adjoint_synthesis_general_spin(r, lmax, loc, spin, 2, ms,
                                            lstride, size=n)
    got_b = at.cpu().numpy()
    used = np.zeros(n, bool)
    for m in range(mmax + 1):
        used[ms[m] + np.arange(m, lmax + 1) * lstride] = True
    assert O.l2error(got_b[:, used].ravel(), bref[:, used].ravel()) <= eps
    assert np.all(got_b[:, ~used] == 7.0)


def _polar_loc(rng, npt):
    loc = _loc(rng, npt)
    k = npt // 3
    loc[:k, 0] = rng.uniform(0, 1e-3, k)
    loc[k:2 * k, 0] = np.pi - rng.uniform(0, 1e-3, k)
    loc[-10:, 0] = np.pi / 2
    return loc


def test_spin_near_poles_vs_direct_sum(gpu):
    """lmax 300, spin 3, points within 1e-3 of the poles and on the
    equator: 200 sampled points against the direct sum"""
    from ducc_amd import sht
    rng = np.random.default_rng(21)
    lmax, spin, eps, npt = 300, 3, 1e-10, 20_000
    a = _random_alm(lmax, lmax, spin, 2, rng)
    loc = _polar_loc(rng, npt)
    got = sht.synthesis_general(alm=a, spin=spin, lmax=lmax, loc=loc,
                                epsilon=eps)
    sel = np.concatenate([rng.choice(npt, 190, replace=False),
                          np.arange(npt - 10, npt)])
    ref = S.synthesis_general_spin(a, lmax, loc[sel], spin)
    assert O.l2error(got[:, sel].ravel().astype(complex),
                     ref.ravel().astype(complex)) <= eps


def test_spin_high_m_vs_ducc0(gpu):
    """lmax 1500, spin 2, a third of the points within 1e-3 of each pole:
    the scaled recursions keep the large-m terms (tiny there) from
    underflowing to garbage; synthesis and adjoint within 2 eps of ducc0's
    own (oracle/_ref)"""
    import os
    from oracle import ref as R
    from ducc_amd import sht
    if not R.available():
        pytest.skip("reference library not built")
    rng = np.random.default_rng(22)
    lmax, spin, eps, npt = 1500, 2, 1e-9, 20_000
    a = _random_alm(lmax, lmax, spin, 2, rng)
    loc = _polar_loc(rng, npt)
    nth = os.cpu_count() or 1
    got = sht.synthesis_general(alm=a, spin=spin, lmax=lmax, loc=loc,
                                epsilon=eps)
    ref = R.sht_synthesis_general(a, lmax, loc, eps, nthreads=nth, spin=spin)
    assert O.l2error(got.ravel().astype(complex),
                     ref.ravel().astype(complex)) <= 2 * eps
    r = rng.uniform(-1, 1, (2, npt))
    got_a = sht.adjoint_synthesis_general(map=r, spin=spin, lmax=lmax,
                                          loc=loc, epsilon=eps)
    ref_a = R.sht_adjoint_synthesis_general(r, lmax, loc, eps, a.shape[1],
                                            nthreads=nth, spin=spin)
    assert O.l2error(got_a.ravel(), ref_a.ravel()) <= 2 * eps


def test_spin_errors(gpu):
    from ducc_amd import sht
    rng = np.random.default_rng(1)
    loc = _loc(rng, 10)
    a2 = _random_alm(4, 4, 2, 2, rng)
    with pytest.raises(RuntimeError, match="number of components"):
        sht.synthesis_general(alm=a2[:1], spin=2, lmax=4, loc=loc,
                              epsilon=1e-5)
    with pytest.raises(RuntimeError, match="number of components"):
        sht.synthesis_general(alm=a2, spin=2, lmax=4, loc=loc, epsilon=1e-5,
                              mode="GRAD_ONLY")
    with pytest.raises(RuntimeError, match="spin>0"):
        sht.synthesis_general(alm=a2[:1], spin=0, lmax=4, loc=loc,
                              epsilon=1e-5, mode="GRAD_ONLY")
    with pytest.raises(RuntimeError, match="spin 1"):
        sht.synthesis_general(alm=a2[:1], spin=2, lmax=4, loc=loc,
                              epsilon=1e-5, mode="DERIV1")
    with pytest.raises(RuntimeError, match="unknown SHT mode"):
        sht.synthesis_general(alm=a2, spin=2, lmax=4, loc=loc, epsilon=1e-5,
                              mode="CURL")
    with pytest.raises(RuntimeError, match="number of components"):
        sht.adjoint_synthesis_general(map=np.zeros((1, 10)), spin=2, lmax=4,
                                      loc=loc, epsilon=1e-5)


@pytest.mark.parametrize("mode", ["STANDARD", "GRAD_ONLY"])
def test_spin_matches_ducc0(gpu, mode):
    """lmax 400, spin 2, 10^5 points, eps 1e-8: synthesis and adjoint
    within 2 eps of ducc0's own (oracle/_ref, all host cores)"""
    import os
    from oracle import ref as R
    from ducc_amd import sht
    if not R.available():
        pytest.skip("reference library not built")
    rng = np.random.default_rng(400)
    lmax, npt, spin, eps = 400, 100_000, 2, 1e-8
    nc = 2 if mode == "STANDARD" else 1
    a = _random_alm(lmax, lmax, spin, nc, rng)
    loc = _loc(rng, npt)
    nth = os.cpu_count() or 1
    got = sht.synthesis_general(alm=a, spin=spin, lmax=lmax, loc=loc,
                                epsilon=eps, mode=mode)
    ref = R.sht_synthesis_general(a, lmax, loc, eps, nthreads=nth, spin=spin,
                                  mode=mode)
    assert O.l2error(got.ravel().astype(complex),
                     ref.ravel().astype(complex)) <= 2 * eps
    r = rng.uniform(-1, 1, (2, npt))
    got_a = sht.adjoint_synthesis_general(map=r, spin=spin, lmax=lmax,
                                          loc=loc, epsilon=eps, mode=mode)
    ref_a = R.sht_adjoint_synthesis_general(r, lmax, loc, eps, a.shape[1],
                                            nthreads=nth, spin=spin,
                                            mode=mode)
    assert O.l2error(got_a.ravel(), ref_a.ravel()) <= 2 * eps
