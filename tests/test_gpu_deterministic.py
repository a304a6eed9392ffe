"""Deterministic mode: type-1 results identical bit for bit from run to run
(no floating-point atomics in the spreading: per-subproblem FP64 regions
summed per grid cell in a fixed order), still within the reference tests'
bounds of the direct sum, and close to the default (atomic) spreader."""
import numpy as np
import pytest
import torch

from oracle import core as O

pytestmark = pytest.mark.gpu


def _inputs(seed, M, shape, clustered=False):
    rng = np.random.default_rng(seed)
    nd = len(shape)
    if clustered:
        cen = rng.uniform(-np.pi, np.pi, (4, nd))
        c = cen[rng.integers(0, 4, M)] + rng.normal(0, 0.2, (M, nd))
        c = (c + np.pi) % (2 * np.pi) - np.pi
    else:
        c = rng.uniform(-np.pi, np.pi, (M, nd))
    p = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    return c, p


@pytest.mark.parametrize("shape,M,eps,dt", [
    ((300,), 50_000, 1e-9, np.complex128),
    ((17,), 20_000, 1e-5, np.complex64),      # regions wrap many times
    ((128, 96), 200_000, 1e-5, np.complex64),
    ((120, 100), 150_000, 1e-11, np.complex128),
    ((9, 7), 30_000, 1e-6, np.complex128),    # tiny grid
    ((24, 20, 16), 100_000, 1e-4, np.complex64),
    ((20, 18, 16), 60_000, 1e-9, np.complex128),
])
@pytest.mark.parametrize("clustered", [False, True])
def test_bitwise_reproducible_and_accurate(gpu, shape, M, eps, dt,
                                           clustered):
    from ducc_amd import nufft
    c, p = _inputs(len(shape) * 100 + M % 97, M, shape, clustered)
    p = p.astype(dt)
    ct = torch.from_numpy(c).cuda()
    pt = torch.from_numpy(p).cuda()
    runs = []
    for k in range(3):
        out = torch.empty(shape, dtype=pt.dtype, device="cuda")
        if k == 1:
            # other work on the device in between, and a side stream
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                nufft.nu2u(points=pt, coord=ct, forward=True, epsilon=eps,
                           out=out, deterministic=True)
            side.synchronize()
        else:
            nufft.nu2u(points=pt, coord=ct, forward=True, epsilon=eps,
                       out=out, deterministic=True)
        runs.append(out.cpu().numpy())
    assert np.array_equal(runs[0].view(np.uint8), runs[1].view(np.uint8))
    assert np.array_equal(runs[0].view(np.uint8), runs[2].view(np.uint8))
    ref_default = torch.empty(shape, dtype=pt.dtype, device="cuda")
    nufft.nu2u(points=pt, coord=ct, forward=True, epsilon=eps,
               out=ref_default)
    d = ref_default.cpu().numpy()
    # both spreaders sum the same FP64 terms in different orders
    tol = 1e-12 if dt == np.complex128 else 2e-6
    assert O.l2error(runs[0], d) <= tol
    mul = {1: 4, 2: 2, 3: 1}[len(shape)]
    N = int(np.prod(shape))
    sel = np.random.default_rng(5).choice(N, min(N, 300), replace=False)
    ref = O.nudft_type1(c, p.astype(np.complex128), shape, True, modes=sel)
    floor = 0 if dt == np.complex128 else 1e-6
    assert O.l2error(runs[0].ravel()[sel], ref) <= mul * max(eps, floor)


def test_deterministic_plan_matches_unplanned(gpu):
    from ducc_amd import nufft
    c, p = _inputs(9, 120_000, (100, 80))
    ct = torch.from_numpy(c).cuda()
    pt = torch.from_numpy(p).cuda()
    a = torch.empty((100, 80), dtype=torch.complex128, device="cuda")
    nufft.nu2u(points=pt, coord=ct, forward=False, epsilon=1e-10, out=a,
               deterministic=True, sigma_min=1.1, sigma_max=2.6)
    pl = nufft.plan(nu2u=True, coord=ct, grid_shape=(100, 80),
                    epsilon=1e-10, deterministic=True)
    b = pl.nu2u(points=pt, forward=False)
    b2 = pl.nu2u(points=pt, forward=False)
    torch.cuda.synchronize()
    assert torch.equal(b, b2)
    assert O.l2error(a.cpu().numpy(), b.cpu().numpy()) < 1e-12
