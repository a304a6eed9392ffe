"""The fold in spreading and interpolation agrees with the fold of the index
build: with check_offsets=True (DMX_CHECK_OFFSETS) every point's offset
inside the tile its key names must lie in [0, tile); the kernels count the
points they would otherwise clamp, and the call fails if any.  Covered: 1-3
D, uniform and clustered points, coordinates on grid points, at +-P/2,
far outside the period and one ulp inside it, float32 coordinates, FP32 and
FP64 compute, types 1 and 2, the point-records path, the deterministic
spreader and plans; results equal those of the unchecked call."""
import numpy as np
import pytest

from oracle import core as O

pytestmark = pytest.mark.gpu


def _edge_coords(ndim, shape, rng, M):
    """random points plus grid points of several oversampled lengths, the
    period's ends and points far outside it"""
    c = rng.uniform(-np.pi, np.pi, (M, ndim))
    j = np.arange(-40, 40)
    for n in (shape[0], 2 * shape[0], 3388, 2880, 616, 594):
        g = 2 * np.pi * j / n
        c = np.concatenate([c, np.repeat(g[:, None], ndim, axis=1)])
    extra = np.array([np.pi, -np.pi, 0.0, 1e3, 7 * np.pi,
                      np.nextafter(np.pi, 0), np.nextafter(-np.pi, 0)])
    c = np.concatenate([c, np.repeat(extra[:, None], ndim, axis=1)])
    return c


CASES = [
    # ndim, shape, M, eps, value dtype, coordinate dtype
    (1, (1000,), 20000, 1e-5, np.complex64, np.float64),
    (2, (256, 200), 50000, 1e-5, np.complex64, np.float64),
    (2, (128, 128), 30000, 1e-12, np.complex128, np.float64),
    (2, (300, 300), 40000, 1e-4, np.complex64, np.float32),
    (3, (48, 40, 32), 30000, 1e-4, np.complex64, np.float64),
    (3, (32, 32, 32), 20000, 1e-9, np.complex128, np.float64),
]


def _ids(c):
    return f"{c[0]}d-{'x'.join(map(str, c[1]))}-{c[3]:g}-" \
           f"{np.dtype(c[4]).name}-{np.dtype(c[5]).name}"


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_no_offset_is_clamped(gpu, case):
    from ducc_amd import nufft
    ndim, shape, M, eps, vdt, cdt = case
    print("offsets", _ids(case), flush=True)
    rng = np.random.default_rng(ndim * 1000 + M)
    c = _edge_coords(ndim, shape, rng, M).astype(cdt)
    n = len(c)
    p = (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(vdt)
    g = (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)) \
        .astype(vdt)
    for kw in ({}, {"point_records": True}, {"deterministic": True}):
        o1 = np.empty(shape, vdt)
        nufft.nu2u(points=p, coord=c, forward=True, epsilon=eps, out=o1,
                   check_offsets=True, **kw)
        o0 = np.empty(shape, vdt)
        nufft.nu2u(points=p, coord=c, forward=True, epsilon=eps, out=o0,
                   **kw)
        assert O.l2error(o1.astype(np.complex128),
                         o0.astype(np.complex128)) <= 1e-6
        if kw.get("deterministic"):
            continue
        r1 = nufft.u2nu(grid=g, coord=c, forward=False, epsilon=eps,
                        check_offsets=True, **kw)
        r0 = nufft.u2nu(grid=g, coord=c, forward=False, epsilon=eps, **kw)
        assert np.array_equal(r1, r0)


def test_plan_checks_offsets(gpu):
    from ducc_amd import nufft
    rng = np.random.default_rng(3)
    c = _edge_coords(2, (512, 512), rng, 100000)
    n = len(c)
    p = (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n))
    for nu2u in (True, False):
        pl = nufft.plan(nu2u=nu2u, coord=c, grid_shape=(512, 512),
                        epsilon=1e-9, check_offsets=True)
        if nu2u:
            out = np.empty((512, 512), np.complex128)
            pl.nu2u(points=p, forward=True, out=out)
            pl.nu2u(points=p, forward=False, out=out)
        else:
            g = rng.uniform(-1, 1, (512, 512)) + 0j
            pl.u2nu(grid=g, forward=True)
            pl.u2nu(grid=g, forward=False)
        del pl


def test_noalloc_refuses_the_check(gpu):
    import ctypes as ct
    import torch
    from ducc_amd import _lib
    c = torch.zeros((10, 2), dtype=torch.float64, device="cuda")
    p = torch.zeros(10, dtype=torch.complex128, device="cuda")
    g = torch.zeros((16, 16), dtype=torch.complex128, device="cuda")
    (dg, _g), (dc, _c), (dp, _p) = (_lib.describe(x) for x in (g, c, p))
    opt = _lib.options(stream=torch.cuda.current_stream().cuda_stream,
                       check_offsets=True)
    per = np.array([2 * np.pi] * 2)
    rc = _lib.lib().dmx_nufft_noalloc(1, ct.byref(dg), ct.byref(dc),
                                      ct.byref(dp), 1, ct.c_double(1e-6),
                                      ct.c_double(1.2), ct.c_double(2.51),
                                      per.ctypes.data, 0, ct.byref(opt), None)
    assert rc != 0
    assert "offset checks" in _lib.lib().dmx_last_error().decode()
