"""ducc0.sht.synthesis_general / adjoint_synthesis_general (spin 0) on the
GPU against the oracle's direct sums (oracle/sht.py, itself pinned to
scipy's spherical harmonics), and the reference's own adjointness test
(python/test/test_sht.py:238-255: random a_lm and points, rtol 1e-9) and
accuracy bound (test_sht.py:268-279: l2error within epsilon)."""
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


CASES = [
    # lmax, mmax, npoints, eps, complex dtype
    (0, 0, 50, 1e-10, np.complex128),
    (5, 5, 300, 1e-10, np.complex128),
    (10, 8, 1000, 1e-7, np.complex128),
    (32, 32, 2000, 1e-12, np.complex128),
    (100, 100, 2000, 1e-5, np.complex128),
    (64, 40, 3000, 1e-4, np.complex64),
    (200, 200, 500, 1e-9, np.complex128),
]


@pytest.mark.parametrize("lmax,mmax,npt,eps,cdt", CASES)
def test_synthesis_vs_direct_sum(gpu, lmax, mmax, npt, eps, cdt):
    from ducc_amd import sht
    print(f"synthesis lmax={lmax} mmax={mmax} npt={npt} eps={eps}",
          flush=True)
    rng = np.random.default_rng(lmax + npt)
    a = S.random_alm(lmax, mmax, rng).astype(cdt)
    loc = _loc(rng, npt)
    got = sht.synthesis_general(alm=a[None, :], spin=0, lmax=lmax, mmax=mmax,
                                loc=loc, epsilon=eps)
    assert got.shape == (1, npt)
    assert got.dtype == (np.float64 if cdt == np.complex128 else np.float32)
    ref = S.synthesis_general(a.astype(np.complex128), lmax, loc,
                              S.default_mstart(lmax, mmax))
    assert O.l2error(got[0].astype(np.complex128), ref) <= eps


@pytest.mark.parametrize("lmax,mmax,npt,eps,cdt", CASES)
def test_adjoint_vs_direct_sum(gpu, lmax, mmax, npt, eps, cdt):
    from ducc_amd import sht
    print(f"adjoint lmax={lmax} mmax={mmax} npt={npt} eps={eps}", flush=True)
    rng = np.random.default_rng(7 + lmax + npt)
    loc = _loc(rng, npt)
    rdt = np.float64 if cdt == np.complex128 else np.float32
    r = rng.uniform(-0.5, 0.5, npt).astype(rdt)
    got = sht.adjoint_synthesis_general(map=r[None, :], spin=0, lmax=lmax,
                                        mmax=mmax, loc=loc, epsilon=eps)
    assert got.dtype == cdt and got.shape == (1, S.nalm(lmax, mmax))
    ref = S.adjoint_synthesis_general(r, lmax, loc,
                                      S.default_mstart(lmax, mmax))
    assert O.l2error(got[0].astype(np.complex128), ref) <= eps


@pytest.mark.parametrize("lmmax", [(10, 8), (100, 100)])
@pytest.mark.parametrize("npix", [33, 200])
def test_adjointness_general(gpu, lmmax, npix):
    """test_sht.py:238-255 for spin 0"""
    from ducc_amd import sht
    rng = np.random.default_rng(48)
    lmax, mmax = lmmax
    eps = 1e-5
    a = S.random_alm(lmax, mmax, rng)[None, :]
    loc = _loc(rng, npix)
    r = rng.uniform(-0.5, 0.5, (1, npix))
    p1 = sht.synthesis_general(lmax=lmax, mmax=mmax, alm=a, loc=loc, spin=0,
                               epsilon=eps)
    a2 = sht.adjoint_synthesis_general(lmax=lmax, mmax=mmax, map=r, loc=loc,
                                       spin=0, epsilon=eps)
    ms = S.default_mstart(lmax, mmax)
    v1 = S.almdot(a[0], a2[0], lmax, ms)
    v2 = float(np.dot(r[0], p1[0]))
    assert abs(v1 - v2) <= 1e-9 * abs(v1)


def test_poles_and_seams(gpu):
    """points on both poles, on phi = 0 and next to 2 pi, on the equator"""
    from ducc_amd import sht
    rng = np.random.default_rng(5)
    lmax, eps = 24, 1e-11
    a = S.random_alm(lmax, lmax, rng)
    loc = np.array([[0.0, 0.0], [0.0, 2.0], [np.pi, 0.5], [np.pi, 6.0],
                    [np.pi / 2, 0.0], [1.0, np.nextafter(2 * np.pi, 0)],
                    [np.nextafter(np.pi, 0), 1.0], [1e-12, 3.0]])
    got = sht.synthesis_general(alm=a[None, :], spin=0, lmax=lmax, loc=loc,
                                epsilon=eps)[0]
    ref = S.synthesis_general(a, lmax, loc)
    assert O.l2error(got.astype(np.complex128), ref) <= eps


def test_layout_stride_and_torch(gpu):
    """custom mstart / lstride, torch device arrays (in place, strided map)
    and numpy arrays give the same values"""
    from ducc_amd import sht
    rng = np.random.default_rng(9)
    lmax, mmax, eps, lstride = 30, 20, 1e-10, 2
    ms = S.default_mstart(lmax, mmax) * lstride + 3
    n = int(max(ms[m] + lmax * lstride for m in range(mmax + 1))) + 1
    a = np.zeros(n, np.complex128)
    for m in range(mmax + 1):
        for l in range(m, lmax + 1):
            a[ms[m] + l * lstride] = rng.uniform(-1, 1) + \
                (1j * rng.uniform(-1, 1) if m else 0)
    loc = _loc(rng, 700)
    ref = S.synthesis_general(a, lmax, loc, ms, lstride)
    got = sht.synthesis_general(alm=a[None, :], spin=0, lmax=lmax,
                                mstart=ms, lstride=lstride, loc=loc,
                                epsilon=eps)[0]
    assert O.l2error(got.astype(np.complex128), ref) <= eps
    mp = torch.zeros((1, 2 * len(loc)), dtype=torch.float64, device="cuda")
    out = mp[:, ::2]
    res = sht.synthesis_general(alm=torch.from_numpy(a[None, :]).cuda(),
                                spin=0, lmax=lmax, mstart=ms,
                                lstride=lstride,
                                loc=torch.from_numpy(loc).cuda(), epsilon=eps,
                                map=out)
    assert res is out
    np.testing.assert_allclose(out.cpu().numpy()[0], got, rtol=0, atol=1e-13)
    assert torch.all(mp[:, 1::2] == 0)
    # adjoint into a device array keeps the entries outside the layout
    r = rng.uniform(-1, 1, len(loc))
    at = torch.full((1, n), 7.0 + 0j, dtype=torch.complex128, device="cuda")
    sht.adjoint_synthesis_general(map=torch.from_numpy(r[None, :]).cuda(),
                                  spin=0, lmax=lmax, mstart=ms,
                                  lstride=lstride,
                                  loc=torch.from_numpy(loc).cuda(),
                                  epsilon=eps, alm=at)
    bref = S.adjoint_synthesis_general(r, lmax, loc, ms, lstride, size=n)
    got_b = at.cpu().numpy()[0]
    used = np.zeros(n, bool)
    for m in range(mmax + 1):
        used[ms[m] + np.arange(m, lmax + 1) * lstride] = True
    assert O.l2error(got_b[used], bref[used]) <= eps
    assert np.all(got_b[~used] == 7.0)


def test_larger_transform_sampled(gpu):
    """lmax 1000, 10^5 points, eps 1e-8: 200 sampled points against the
    direct sum"""
    from ducc_amd import sht
    rng = np.random.default_rng(11)
    lmax, eps, npt = 1000, 1e-8, 100_000
    a = S.random_alm(lmax, lmax, rng)
    loc = _loc(rng, npt)
    got = sht.synthesis_general(alm=a[None, :], spin=0, lmax=lmax, loc=loc,
                                epsilon=eps)[0]
    sel = rng.choice(npt, 200, replace=False)
    ref = S.synthesis_general(a, lmax, loc[sel])
    assert O.l2error(got[sel].astype(np.complex128), ref) <= eps


def test_errors(gpu):
    from ducc_amd import sht
    rng = np.random.default_rng(1)
    a = S.random_alm(4, 4, rng)[None, :]
    loc = _loc(rng, 10)
    with pytest.raises(RuntimeError, match="number of components"):
        sht.synthesis_general(alm=a, spin=1, lmax=4, loc=loc, epsilon=1e-5)
    with pytest.raises(RuntimeError, match="loc"):
        sht.synthesis_general(alm=a, spin=0, lmax=4, loc=loc[:, :1].copy(),
                              epsilon=1e-5)
    with pytest.raises(RuntimeError, match="mmax>lmax"):
        sht.synthesis_general(alm=a, spin=0, lmax=4, mmax=5, loc=loc,
                              epsilon=1e-5)


@pytest.mark.parametrize("cdt,eps", [(np.complex128, 1e-8),
                                     (np.complex64, 1e-4)])
def test_matches_ducc0(gpu, cdt, eps):
    """lmax 500, 10^5 points: synthesis and adjoint within 2 eps of ducc0's
    own synthesis_general / adjoint_synthesis_general (oracle/_ref, all host
    cores)"""
    import os
    from oracle import ref as R
    from ducc_amd import sht
    if not R.available():
        pytest.skip("reference library not built")
    rng = np.random.default_rng(500)
    lmax, npt = 500, 100_000
    a = S.random_alm(lmax, lmax, rng).astype(cdt)
    loc = _loc(rng, npt)
    nth = os.cpu_count() or 1
    got = sht.synthesis_general(alm=a[None, :], spin=0, lmax=lmax, loc=loc,
                                epsilon=eps)[0]
    ref = R.sht_synthesis_general(a, lmax, loc, eps, nthreads=nth)
    assert O.l2error(got.astype(np.complex128),
                     ref.astype(np.complex128)) <= 2 * eps
    rdt = np.float64 if cdt == np.complex128 else np.float32
    r = rng.uniform(-1, 1, npt).astype(rdt)
    got_a = sht.adjoint_synthesis_general(map=r[None, :], spin=0, lmax=lmax,
                                          loc=loc, epsilon=eps)[0]
    ref_a = R.sht_adjoint_synthesis_general(r, lmax, loc, eps, len(a),
                                            nthreads=nth)
    assert O.l2error(got_a.astype(np.complex128),
                     ref_a.astype(np.complex128)) <= 2 * eps
