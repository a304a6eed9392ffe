"""The record sort of unplanned transforms (the points themselves sorted into
tile order, carrying their tile-local offsets, kernel arguments and value or
input index) against the oracle, bit for bit where the arithmetic is
integer: the order is the stable tile order of the index build
(nufft.h:1364-1381, bucket_sort.h:170-354 ordering), the offsets are the
fold's i0 (nufft.h:285-298) within the tile.  Transforms on the record path
(DMX_POINT_RECORDS) are compared with the default index path on the same
inputs: type
2 reads the same cells with the same weights in the same order, so it is
identical bit for bit; type 1 differs only in the order of its floating-point
atomics."""
import numpy as np
import pytest
import torch

from oracle import core as O

pytestmark = pytest.mark.gpu


def _clustered(rng, M, d, ncl=8, sig=0.3):
    cen = rng.uniform(-np.pi, np.pi, (ncl, d))
    lab = rng.integers(0, ncl, M)
    x = cen[lab] + rng.normal(0, sig, (M, d))
    return (x + np.pi) % (2 * np.pi) - np.pi


SORT_CASES = [
    (1, (4096,), 9, 7, 100000, "uniform"),
    (2, (3388, 3388), 5, 7, 300000, "cluster"),
    (2, (462, 462), 4, 11, 50001, "uniform"),
    (2, (40, 16), 5, 8, 1000, "uniform"),
    (3, (432, 432, 432), 3, 6, 200000, "uniform"),
    (3, (64, 80, 96), 2, 16, 20000, "cluster"),
    (2, (3388, 3388), 5, 7, 1, "uniform"),
]


@pytest.mark.parametrize("ndim,nover,log2tile,W,M,dist", SORT_CASES)
def test_record_sort_bit_exact(gpu, ndim, nover, log2tile, W, M, dist):
    from ducc_amd import nufft
    rng = np.random.default_rng(2000 + M)
    if dist == "uniform":
        c = rng.uniform(-np.pi, np.pi, (M, ndim))
    else:
        c = _clustered(rng, M, ndim)
    words, idx, xarg = nufft.sort_records(c, nover, W, log2tile)
    kref = O.tile_keys(c, W, nover, 2 * np.pi, log2tile)
    iref = np.argsort(kref, kind="stable").astype(np.uint32)
    assert np.array_equal(idx, iref)
    ob = ndim * log2tile
    assert np.array_equal(words >> np.uint32(ob), kref[iref])
    i0, frac = O.fold(c, W, nover, 2 * np.pi)
    nsafe = (W + 1) // 2
    tm = (1 << log2tile) - 1
    for d in range(ndim):
        off = (words >> np.uint32(d * log2tile)) & np.uint32(tm)
        want = (i0[iref, d] + nsafe) & tm
        assert np.array_equal(off, want.astype(np.uint32))
        # the kernel argument from the GPU's double-double fold and the
        # oracle's 80-bit one: equal to a few units of the last place
        x_ref = (W - 1) - 2.0 * frac[iref, d]
        assert np.max(np.abs(xarg[:, d] - x_ref)) <= 1e-12


def test_record_sort_edge_points(gpu):
    """Coordinates on grid points, at +-P/2, 0 and far outside [-P/2, P/2)
    (as for the index sort)."""
    from ducc_amd import nufft
    nover = (3388, 2880)
    j = np.arange(-50, 50)
    c = np.stack([2 * np.pi * j / n for n in nover], axis=1)
    extra = np.array([[np.pi, -np.pi], [0.0, 0.0], [-np.pi, np.pi],
                      [1e3, -1e3], [7 * np.pi, -5 * np.pi],
                      [np.nextafter(np.pi, 0), np.nextafter(-np.pi, 0)]])
    c = np.concatenate([c, extra])
    for W, l2 in ((7, 5), (8, 4), (16, 4), (4, 5)):
        words, idx, _ = nufft.sort_records(c, nover, W, l2)
        kref = O.tile_keys(c, W, nover, 2 * np.pi, l2)
        iref = np.argsort(kref, kind="stable")
        assert np.array_equal(idx, iref)
        assert np.array_equal(words >> np.uint32(2 * l2), kref[iref])


def _inputs(seed, nd, M, shape, cdt, clustered):
    rng = np.random.default_rng(seed)
    c = (_clustered(rng, M, nd) if clustered
         else rng.uniform(-np.pi, np.pi, (M, nd)))
    p = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    g = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return (c, torch.from_numpy(c).cuda(), torch.from_numpy(p).cuda().to(cdt),
            torch.from_numpy(g).cuda().to(cdt))


TRANSFORM_CASES = [
    (1, 200_000, (5000,), 1e-5, torch.complex64, False),
    (1, 100_000, (3000,), 1e-10, torch.complex128, False),
    (2, 1_500_000, (512, 384), 1e-5, torch.complex64, True),
    (2, 300_000, (300, 200), 1e-11, torch.complex128, False),
    (3, 400_000, (64, 48, 40), 1e-4, torch.complex64, False),
    (3, 100_000, (32, 40, 24), 1e-9, torch.complex128, True),
    (2, 4097, (64, 64), 1e-6, torch.complex64, False),
]


@pytest.mark.parametrize("case", range(len(TRANSFORM_CASES)))
def test_u2nu_records_equal_index_path(gpu, case):
    from ducc_amd import nufft
    nd, M, shape, eps, cdt, cl = TRANSFORM_CASES[case]
    print(f"u2nu records vs index: {nd}-D M={M} {shape} eps={eps} {cdt}",
          flush=True)
    c, ct_, pt, gt = _inputs(50 + case, nd, M, shape, cdt, cl)
    i1, i2 = {}, {}
    a = nufft.u2nu(grid=gt, coord=ct_, forward=False, epsilon=eps,
                   point_records=True, info=i1)
    b = nufft.u2nu(grid=gt, coord=ct_, forward=False, epsilon=eps, info=i2)
    assert i1 == i2
    assert torch.equal(a, b)
    sel = np.random.default_rng(case).choice(M, 64, replace=False)
    ref = O.nudft_type2(c[sel], gt.cpu().numpy().astype(np.complex128), False)
    mul = {1: 4, 2: 2, 3: 1}[nd]
    assert O.l2error(a.cpu().numpy()[sel], ref) <= mul * max(eps, 2e-6)


@pytest.mark.parametrize("case", range(len(TRANSFORM_CASES)))
def test_nu2u_records_match_index_path(gpu, case):
    from ducc_amd import nufft
    nd, M, shape, eps, cdt, cl = TRANSFORM_CASES[case]
    print(f"nu2u records vs index: {nd}-D M={M} {shape} eps={eps} {cdt}",
          flush=True)
    c, ct_, pt, gt = _inputs(60 + case, nd, M, shape, cdt, cl)
    a = torch.empty(shape, dtype=cdt, device="cuda")
    b = torch.empty(shape, dtype=cdt, device="cuda")
    nufft.nu2u(points=pt, coord=ct_, forward=True, epsilon=eps, out=a,
               point_records=True)
    nufft.nu2u(points=pt, coord=ct_, forward=True, epsilon=eps, out=b)
    # same sums, in another order of the floating-point atomics
    tol = 1e-5 if cdt == torch.complex64 else 1e-13
    assert O.l2error(a.cpu().numpy(), b.cpu().numpy()) <= tol
    n = int(np.prod(shape))
    sel = np.random.default_rng(case).choice(n, min(64, n), replace=False)
    ref = O.nudft_type1(c, pt.cpu().numpy().astype(np.complex128), shape,
                        True, modes=sel)
    mul = {1: 4, 2: 2, 3: 1}[nd]
    assert O.l2error(a.cpu().numpy().ravel()[sel], ref) <= mul * max(eps,
                                                                      2e-6)


def test_host_arrays_take_the_record_path(gpu):
    """numpy inputs (staged by the library) on the record path give the
    same type-2 values as torch inputs on the index path, and type 1
    agrees to rounding."""
    from ducc_amd import nufft
    c, ct_, pt, gt = _inputs(7, 2, 100_000, (128, 96), torch.complex64, True)
    a = nufft.u2nu(grid=gt.cpu().numpy(), coord=c, forward=True,
                   epsilon=1e-5, point_records=True)
    b = nufft.u2nu(grid=gt, coord=ct_, forward=True, epsilon=1e-5)
    assert np.array_equal(a, b.cpu().numpy())
    o = np.empty((128, 96), np.complex64)
    nufft.nu2u(points=pt.cpu().numpy(), coord=c, forward=False,
               epsilon=1e-5, out=o, point_records=True)
    o2 = torch.empty((128, 96), dtype=torch.complex64, device="cuda")
    nufft.nu2u(points=pt, coord=ct_, forward=False, epsilon=1e-5, out=o2)
    assert O.l2error(o, o2.cpu().numpy()) <= 1e-5
