"""Bit-exact parity of the GPU index build (fold + tile key + stable sort)
against the CPU oracle (nufft.h:285-320, 892-907, 1364-1381;
bucket_sort.h:170-354 ordering)."""
import numpy as np
import pytest

from oracle import core as O

pytestmark = pytest.mark.gpu


def _clustered(rng, M, d, ncl=8, sig=0.3):
    cen = rng.uniform(-np.pi, np.pi, (ncl, d))
    lab = rng.integers(0, ncl, M)
    x = cen[lab] + rng.normal(0, sig, (M, d))
    return (x + np.pi) % (2 * np.pi) - np.pi


CASES = [
    (1, (4096,), 9, 7, 100000, "uniform"),
    (2, (3388, 3388), 5, 7, 300000, "cluster"),
    (2, (462, 462), 4, 11, 50000, "uniform"),
    (2, (40, 16), 5, 8, 1000, "uniform"),
    (3, (432, 432, 432), 3, 6, 200000, "uniform"),
    (3, (64, 80, 96), 2, 16, 20000, "cluster"),
]


@pytest.mark.parametrize("ndim,nover,log2tile,W,M,dist", CASES)
def test_sort_bit_exact(gpu, ndim, nover, log2tile, W, M, dist):
    from ducc_amd import nufft
    rng = np.random.default_rng(1000 + M)
    if dist == "uniform":
        c = rng.uniform(-np.pi, np.pi, (M, ndim))
    else:
        c = _clustered(rng, M, ndim)
    keys, idx = nufft.sort_points(c, nover, W, log2tile)
    kref = O.tile_keys(c, W, nover, 2 * np.pi, log2tile)
    assert np.array_equal(keys, kref)
    iref = np.argsort(kref, kind="stable").astype(np.uint32)
    assert np.array_equal(idx, iref)
    assert np.array_equal(idx, O.sort_stable(kref))


def test_sort_edge_points(gpu):
    """Coordinates on grid points, at +-P/2, at 0 and far outside [-P/2,P/2):
    the fold must give the oracle's tile for every one."""
    from ducc_amd import nufft
    nover = (3388, 2880)
    j = np.arange(-50, 50)
    pts = []
    for n in nover:
        pts.append(2 * np.pi * j / n)
    c = np.stack(pts, axis=1)
    extra = np.array([[np.pi, -np.pi], [0.0, 0.0], [-np.pi, np.pi],
                      [1e3, -1e3], [7 * np.pi, -5 * np.pi],
                      [np.nextafter(np.pi, 0), np.nextafter(-np.pi, 0)]])
    c = np.concatenate([c, extra])
    for W, l2 in ((7, 5), (8, 4), (16, 4), (4, 5)):
        keys, idx = nufft.sort_points(c, nover, W, l2)
        kref = O.tile_keys(c, W, nover, 2 * np.pi, l2)
        assert np.array_equal(keys, kref)
        assert np.array_equal(idx, np.argsort(kref, kind="stable"))


def test_sort_float32_coords(gpu):
    from ducc_amd import nufft
    rng = np.random.default_rng(5)
    c = rng.uniform(-np.pi, np.pi, (100000, 2)).astype(np.float32)
    keys, idx = nufft.sort_points(c, (1210, 1210), 7, 5)
    kref = O.tile_keys(c.astype(np.float64), 7, (1210, 1210), 2 * np.pi, 5,
                       coord_double=False)
    assert np.array_equal(keys, kref)
    assert np.array_equal(idx, np.argsort(kref, kind="stable"))


def test_sort_single_point(gpu):
    from ducc_amd import nufft
    c = np.array([[0.3, -1.2]])
    keys, idx = nufft.sort_points(c, (64, 64), 6, 4)
    assert idx.tolist() == [0]
    assert keys.tolist() == O.tile_keys(c, 6, (64, 64), 2 * np.pi, 4).tolist()
