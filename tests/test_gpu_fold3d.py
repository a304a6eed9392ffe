"""3-D type 1 with stored tile regions folded into the grid (no grid
zeroing, no global atomics; spread_impl.h slab_out, spread_det.hip
k_fold_regions): FP32 with fixed-point and with FP64 LDS accumulation and
FP64, points that put thousands into one tile (later subproblems of a tile
add to its stored region), grids small enough for the general pair walk
of the fold; against the FP64 direct sum (nufft.nudft_direct, pinned to the
oracle's NUDFT) and against the deterministic mode (its own FP64 slabs and
fold); a NaN value gives an all-NaN result."""
import numpy as np
import pytest

from oracle import core as O

pytestmark = pytest.mark.gpu


def _points(rng, M, dist):
    if dist == "hot":
        c = rng.normal(0.4, 0.01, (M, 3))
    elif dist == "cluster":
        cen = rng.uniform(-np.pi, np.pi, (6, 3))
        c = cen[rng.integers(0, 6, M)] + rng.normal(0, 0.2, (M, 3))
    else:
        c = rng.uniform(-np.pi, np.pi, (M, 3))
    return (c + np.pi) % (2 * np.pi) - np.pi


CASES = [
    # shape, M, eps, points, values
    ((40, 36, 32), 300_000, 1e-6, "hot", np.complex64),
    ((64, 48, 40), 400_000, 1e-5, "cluster", np.complex64),
    ((32, 32, 32), 200_000, 1e-9, "hot", np.complex128),
    ((24, 20, 16), 100_000, 1e-4, "uniform", np.complex64),
    ((9, 12, 10), 20_000, 1e-3, "uniform", np.complex64),
]


def _ids(c):
    return f"{'x'.join(map(str, c[0]))}-M{c[1]}-{c[2]:g}-{c[3]}-" \
           f"{np.dtype(c[4]).name}"


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_region_fold(gpu, case):
    from ducc_amd import nufft
    shape, M, eps, dist, vdt = case
    print("fold3d", _ids(case), flush=True)
    rng = np.random.default_rng(M + shape[0])
    c = _points(rng, M, dist)
    v = (rng.normal(size=M) + 1j * rng.normal(size=M)).astype(vdt)
    out = np.empty(shape, vdt)
    nufft.nu2u(points=v, coord=c, forward=True, epsilon=eps, out=out,
               check_offsets=True)
    det = np.empty(shape, vdt)
    nufft.nu2u(points=v, coord=c, forward=True, epsilon=eps, out=det,
               deterministic=True)
    n = int(np.prod(shape))
    sel = rng.choice(n, min(n, 400), replace=False)
    d = nufft.nudft_direct(1, c, v.astype(np.complex128), shape, True, sel)
    err = O.l2error(out.ravel()[sel].astype(np.complex128), d)
    diff = O.l2error(out.astype(np.complex128), det.astype(np.complex128))
    print(f"err {err:.2e}, vs deterministic {diff:.2e}")
    assert err <= eps
    # the same kernel and grid, only the accumulation differs (FP32: float
    # regions summed in float, a few roundings of 6e-8 per cell)
    bound = 1e-3 * eps if vdt == np.complex128 else max(0.1 * eps, 5e-7)
    assert diff <= bound


@pytest.mark.parametrize("eps", [1e-4, 1e-6])
def test_region_fold_non_finite(gpu, eps):
    """a NaN value turns the whole result into NaN (with fixed point the
    subproblem's region is stored as NaN, otherwise the FP64 sums carry
    it)"""
    from ducc_amd import nufft
    rng = np.random.default_rng(8)
    c = rng.uniform(-np.pi, np.pi, (30_000, 3))
    v = (rng.normal(size=30_000) + 1j * rng.normal(size=30_000)) \
        .astype(np.complex64)
    v[777] = np.nan
    out = np.empty((32, 32, 32), np.complex64)
    nufft.nu2u(points=v, coord=c, forward=True, epsilon=eps, out=out)
    assert np.all(np.isnan(out))
