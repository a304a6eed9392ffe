"""Host-side checks of the c2c drop-in: the numpy.fft checker pinned by the
direct DFT sum, and the argument handling of ducc0.fft.c2c
(python/fft_pymod.cc:73-94, 104-120) without a GPU."""
import numpy as np
import pytest

from ducc_amd import fft


def l2error(a, b):
    a = np.asarray(a).astype(np.complex128)
    b = np.asarray(b).astype(np.complex128)
    return np.sqrt(np.sum(np.abs(a - b) ** 2) /
                   max(np.sum(np.abs(a) ** 2), np.sum(np.abs(b) ** 2), 1e-300))


def dft(x, axis, sign):
    """direct DFT sum along one axis (FP64), the textbook definition"""
    x = np.moveaxis(np.asarray(x, np.complex128), axis, -1)
    n = x.shape[-1]
    k = np.arange(n)
    m = np.exp(sign * 2j * np.pi * (np.outer(k, k) % n) / n)
    return np.moveaxis(x @ m.T, -1, axis)


def test_numpy_pinned_by_direct_sum():
    rng = np.random.default_rng(1)
    for n in (1, 2, 7, 17, 60, 97, 128, 137):
        x = rng.random(n) - 0.5 + 1j * (rng.random(n) - 0.5)
        assert l2error(np.fft.fft(x), dft(x, 0, -1)) < 1e-14
        assert l2error(np.fft.ifft(x) * n, dft(x, 0, 1)) < 1e-14



def test_numpy_fftn_pinned_by_direct_sum():
    rng = np.random.default_rng(2)
    a = rng.random((6, 7, 5)) - 0.5 + 1j * (rng.random((6, 7, 5)) - 0.5)
    ref = dft(dft(a, 2, -1), 0, -1)
    assert l2error(np.fft.fftn(a, axes=(0, 2)), ref) < 1e-14


def test_makeaxes():
    assert fft._makeaxes(3, None) == [0, 1, 2]
    assert fft._makeaxes(3, [-1, 0]) == [2, 0]
    assert fft._makeaxes(2, 1) == [1]
    with pytest.raises(RuntimeError, match="bad axes argument"):
        fft._makeaxes(2, [])
    with pytest.raises(RuntimeError, match="bad axes argument"):
        fft._makeaxes(2, [0, 1, 0])
    with pytest.raises(ValueError, match="exceeds dimensionality"):
        fft._makeaxes(2, [2])
    with pytest.raises(ValueError, match="exceeds dimensionality"):
        fft._makeaxes(2, [-3])


def test_norm_fct():
    shp = (4, 6, 5)
    assert fft._norm_fct(0, shp, [0, 1]) == 1.0
    assert fft._norm_fct(2, shp, [0, 1]) == 1.0 / 24
    assert abs(fft._norm_fct(1, shp, [2, 2]) - 1 / 5) < 1e-16
    with pytest.raises(ValueError, match="invalid value for inorm"):
        fft._norm_fct(3, shp, [0])


def test_reference_c2c_matches_numpy():
    """ducc0::c2c (oracle/_ref) agrees with the numpy checker"""
    from oracle import ref
    if not ref.available():
        pytest.skip("reference library not built (needs the ducc0 sources)")
    rng = np.random.default_rng(3)
    for shp, ax in (((1999,), None), ((24, 35), (0,)), ((6, 7, 11), None)):
        a = rng.random(shp) - 0.5 + 1j * (rng.random(shp) - 0.5)
        assert l2error(ref.c2c(a, axes=ax), np.fft.fftn(a, axes=ax)) < 2e-15
        assert l2error(ref.c2c(a, axes=ax, forward=False),
                       np.fft.ifftn(a, axes=ax) *
                       np.prod([a.shape[i] for i in
                                (ax if ax else range(a.ndim))])) < 2e-15


def test_good_size():
    """smallest 11-smooth (5-smooth for real) length >= n, by brute force"""
    def smooth(m, ps):
        for p in ps:
            while m % p == 0:
                m //= p
        return m == 1
    for n in list(range(0, 300)) + [1021, 4097, 65537, 1000003]:
        for real, ps in ((False, (2, 3, 5, 7, 11)), (True, (2, 3, 5))):
            g = fft.good_size(n, real)
            assert g >= n
            if n > 0:
                assert smooth(g, ps)
                assert all(not smooth(m, ps) for m in range(n, g))
    with pytest.raises(ValueError):
        fft.good_size(-1)
