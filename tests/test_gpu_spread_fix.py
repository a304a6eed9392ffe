"""FP32 type-1 spreading with the tile region in packed fixed point (one
64-bit LDS integer atomic per tap, spread_impl.h FIX; the FP32 path's
default at eps >= 1e-5): within eps of the FP64 direct sum
(nufft.nudft_direct, pinned to the oracle's NUDFT) and within a small
fraction of eps of the FP64-accumulating spreader (fp64_accumulation=True)
on the same inputs, in 1-D, 2-D and 3-D, for uniform, clustered and
single-tile ("hot") point sets, wide value ranges, all-zero values,
records and the deterministic mode (which keeps its own spreader)."""
import numpy as np
import pytest

from oracle import core as O

pytestmark = pytest.mark.gpu


def _points(rng, M, nd, dist):
    if dist == "uniform":
        return rng.uniform(-np.pi, np.pi, (M, nd))
    if dist == "hot":
        c = rng.normal(0.3, 0.004, (M, nd))
    else:
        cen = rng.uniform(-np.pi, np.pi, (8, nd))
        c = cen[rng.integers(0, 8, M)] + rng.normal(0, 0.3, (M, nd))
    return (c + np.pi) % (2 * np.pi) - np.pi


CASES = [
    # shape, M, eps, dist, value scale spread (log10)
    ((100000,), 300_000, 1e-5, "cluster", 0),
    ((256, 256), 200_000, 1e-2, "uniform", 0),
    ((300, 200), 200_000, 1e-3, "cluster", 0),
    ((512, 512), 300_000, 1e-4, "cluster", 0),
    ((1024, 1024), 1_000_000, 1e-5, "cluster", 0),
    ((400, 400), 500_000, 1e-5, "hot", 0),
    ((512, 512), 300_000, 1e-5, "cluster", 6),
    ((48, 40, 32), 300_000, 1e-4, "uniform", 0),
    ((64, 64, 64), 500_000, 1e-5, "cluster", 0),
    ((32, 32, 32), 200_000, 1e-3, "hot", 0),
]


def _ids(c):
    return f"{'x'.join(map(str, c[0]))}-M{c[1]}-{c[2]:g}-{c[3]}-v{c[4]}"


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_fixed_point_spreading(gpu, case):
    from ducc_amd import nufft
    shape, M, eps, dist, vspread = case
    print("fix", _ids(case), flush=True)
    nd = len(shape)
    rng = np.random.default_rng(M + nd + int(-np.log10(eps) * 10))
    c = _points(rng, M, nd, dist)
    v = rng.normal(size=M) + 1j * rng.normal(size=M)
    if vspread:
        v *= 10.0 ** rng.uniform(-vspread, 0, M)
    v = v.astype(np.complex64)
    info = {}
    out = np.empty(shape, np.complex64)
    nufft.nu2u(points=v, coord=c, forward=True, epsilon=eps, out=out,
               info=info, precision="fp32", check_offsets=True)
    assert not info["fp64"]
    ref = np.empty(shape, np.complex64)
    nufft.nu2u(points=v, coord=c, forward=True, epsilon=eps, out=ref,
               precision="fp32", fp64_accumulation=True)
    n = int(np.prod(shape))
    sel = rng.choice(n, min(n, 512), replace=False)
    d = nufft.nudft_direct(1, c, v.astype(np.complex128), shape, True, sel)
    e_fix = O.l2error(out.ravel()[sel].astype(np.complex128), d)
    e_64 = O.l2error(ref.ravel()[sel].astype(np.complex128), d)
    diff = O.l2error(out.astype(np.complex128), ref.astype(np.complex128))
    print(f"W={info['W']} err fix {e_fix:.2e} fp64-acc {e_64:.2e} "
          f"fix vs fp64-acc {diff:.2e}")
    assert e_fix <= eps
    assert diff <= 0.2 * eps
    recs = np.empty(shape, np.complex64)
    nufft.nu2u(points=v, coord=c, forward=True, epsilon=eps, out=recs,
               precision="fp32", point_records=True)
    assert O.l2error(recs.astype(np.complex128),
                     out.astype(np.complex128)) <= 0.2 * eps


def test_fixed_point_zero_values(gpu):
    """all-zero values (no scale exists) give an all-zero grid"""
    from ducc_amd import nufft
    rng = np.random.default_rng(4)
    c = rng.uniform(-0.05, 0.05, (900, 2))
    z = np.zeros(900, np.complex64)
    out = np.full((256, 256), 1 + 1j, np.complex64)
    nufft.nu2u(points=z, coord=c, forward=True, epsilon=1e-4, out=out)
    assert np.all(out == 0)


def test_fixed_point_non_finite_values(gpu):
    """a NaN or infinite value turns the result into NaN, as FP64
    accumulation does (after the FFT every mode sees the bad cell)"""
    from ducc_amd import nufft
    rng = np.random.default_rng(6)
    c = rng.uniform(-np.pi, np.pi, (5000, 2))
    for bad in (np.nan, np.inf):
        v = (rng.normal(size=5000) + 1j * rng.normal(size=5000)) \
            .astype(np.complex64)
        v[1234] = bad
        out = np.empty((64, 64), np.complex64)
        nufft.nu2u(points=v, coord=c, forward=True, epsilon=1e-4, out=out)
        ref = np.empty((64, 64), np.complex64)
        nufft.nu2u(points=v, coord=c, forward=True, epsilon=1e-4, out=ref,
                   fp64_accumulation=True)
        assert not np.all(np.isfinite(ref))
        assert np.all(np.isnan(out))


@pytest.mark.parametrize("M,dist,eps,tile", [(30_000, "uniform", 1e-5, 6),
                                              (30_000, "cluster", 1e-5, 6),
                                              (400_000, "uniform", 1e-5, 5),
                                              (30_000, "cluster", 1e-6, 6)])
def test_sparse_default_tile(gpu, M, dist, eps, tile):
    """2-D FP32 below one point per sixteen oversampled cells takes 64 x 64
    tiles by default (denser sets 32 x 32), with fixed-point and with FP64
    accumulation: type 1 and type 2 within eps of the FP64 direct sum and
    close to explicit 32 x 32 tiles"""
    from ducc_amd import nufft
    shape = (512, 512)
    print("sparse tile", M, dist, eps, flush=True)
    rng = np.random.default_rng(M + len(dist))
    c = _points(rng, M, 2, dist)
    v = (rng.normal(size=M) + 1j * rng.normal(size=M)).astype(np.complex64)
    g = (rng.normal(size=shape) + 1j * rng.normal(size=shape)) \
        .astype(np.complex64)
    i1, i2 = {}, {}
    o1 = np.empty(shape, np.complex64)
    nufft.nu2u(points=v, coord=c, forward=True, epsilon=eps, out=o1,
               info=i1, check_offsets=True)
    o2 = np.empty(M, np.complex64)
    nufft.u2nu(grid=g, coord=c, forward=False, epsilon=eps, out=o2, info=i2,
               check_offsets=True)
    assert i1["log2tile"] == tile and i2["log2tile"] == tile
    assert 16 * M < np.prod(i1["nover"]) if tile == 6 else True
    r1 = np.empty(shape, np.complex64)
    nufft.nu2u(points=v, coord=c, forward=True, epsilon=eps, out=r1,
               log2tile=5)
    r2 = np.empty(M, np.complex64)
    nufft.u2nu(grid=g, coord=c, forward=False, epsilon=eps, out=r2,
               log2tile=5)
    sm = rng.choice(shape[0] * shape[1], 400, replace=False)
    sp = rng.choice(M, 400, replace=False)
    d1 = nufft.nudft_direct(1, c, v.astype(np.complex128), shape, True, sm)
    d2 = nufft.nudft_direct(2, c, g.astype(np.complex128), shape, False, sp)
    e1 = O.l2error(o1.ravel()[sm].astype(np.complex128), d1)
    e2 = O.l2error(o2[sp].astype(np.complex128), d2)
    t1 = O.l2error(o1.astype(np.complex128), r1.astype(np.complex128))
    t2 = O.l2error(o2.astype(np.complex128), r2.astype(np.complex128))
    print(f"err {e1:.2e} {e2:.2e}, vs 32x32 tiles {t1:.2e} {t2:.2e}")
    assert e1 <= eps and e2 <= eps
    assert t1 <= 0.2 * eps and t2 <= 1e-6
