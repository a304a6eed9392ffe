"""The w-gridder restatement (oracle/wgridder.py) pinned independently:
without w-gridding and without phase centre shifts it is the real part of
a 2-D type-1 NUDFT of the weighted visibilities at (2 pi pixsize u,
2 pi pixsize v), its pair is adjoint (test_wgridder.py:197-235), and with
w-gridding a visibility with w = 0 does not see the w-term."""
import numpy as np
import pytest

from oracle import core as O
from oracle import wgridder as G

C = G.SPEED_OF_LIGHT


def _problem(rng, nrow, nchan, nx, ny, fov_deg, wscale=1.0):
    px = fov_deg * np.pi / 180 / nx
    py = fov_deg * np.pi / 180 / ny * 1.1
    f0 = 1e9
    freq = f0 + np.arange(nchan) * (f0 / nchan)
    uvw = (rng.random((nrow, 3)) - 0.5) / (px * f0 / C)
    uvw[:, 2] *= wscale
    ms = rng.random((nrow, nchan)) - 0.5 + 1j * (rng.random((nrow, nchan)) - 0.5)
    wgt = rng.uniform(0.9, 1.1, (nrow, nchan))
    mask = (rng.uniform(0, 1, (nrow, nchan)) > 0.3).astype(np.uint8)
    return px, py, freq, uvw, ms, wgt, mask


def test_no_wgridding_is_a_2d_nudft():
    rng = np.random.default_rng(1)
    nx, ny = 16, 12
    px, py, freq, uvw, ms, wgt, mask = _problem(rng, 7, 3, nx, ny, 2.0)
    d = G.vis2dirty(uvw, freq, ms, wgt, nx, ny, px, py, False, mask)
    uvwe = uvw[:, None, :] * (freq[None, :, None] / C)
    c = np.stack([2 * np.pi * px * uvwe[..., 0].ravel(),
                  2 * np.pi * py * uvwe[..., 1].ravel()], axis=1)
    v = (ms * wgt * (mask != 0)).ravel()
    ref = O.nudft_type1(c, v, (nx, ny), False).real
    np.testing.assert_allclose(d, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("wg", [False, True])
@pytest.mark.parametrize("flips", [(False, False, False), (True, False, True)])
def test_adjointness(wg, flips):
    rng = np.random.default_rng(2)
    nx, ny = 20, 16
    px, py, freq, uvw, ms, wgt, mask = _problem(rng, 9, 4, nx, ny, 40.0)
    dirty = rng.random((nx, ny)) - 0.5
    kw = dict(flip_u=flips[0], flip_v=flips[1], flip_w=flips[2],
              center_x=0.01, center_y=-0.02)
    d = G.vis2dirty(uvw, freq, ms, wgt, nx, ny, px, py, wg, mask, **kw)
    v = G.dirty2vis(uvw, freq, dirty, wgt, px, py, wg, mask, **kw)
    lhs = np.vdot(ms, v).real
    rhs = np.sum(d * dirty)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1)


def test_w_term_vanishes_at_w_zero():
    rng = np.random.default_rng(3)
    nx, ny = 10, 10
    px, py, freq, uvw, ms, wgt, mask = _problem(rng, 5, 2, nx, ny, 10.0,
                                                wscale=0.0)
    a = G.vis2dirty(uvw, freq, ms, wgt, nx, ny, px, py, True, mask,
                    divide_by_n=False)
    b = G.vis2dirty(uvw, freq, ms, wgt, nx, ny, px, py, False, mask)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("wg", [False, True])
def test_matches_ducc0_itself(wg):
    """the direct sums against ducc0's own ms2dirty / dirty2ms (compiled
    from the reference sources into oracle/_ref) at eps 1e-10"""
    from oracle import ref as R
    if not R.available():
        pytest.skip("reference library not built")
    rng = np.random.default_rng(4)
    nx, ny = 32, 36
    px, py, freq, uvw, ms, wgt, mask = _problem(rng, 20, 3, nx, ny, 20.0)
    kw = dict(flip_v=True, center_x=0.003, center_y=0.001)
    d = G.vis2dirty(uvw, freq, ms, wgt, nx, ny, px, py, wg, mask, **kw)
    r = R.wgridder(False, uvw, freq, ms, wgt, mask, nx, ny, px, py, 1e-10,
                   wg, **kw)
    assert O.l2error(d.astype(complex), r.astype(complex)) <= 1e-10
    dirty = rng.random((nx, ny)) - 0.5
    v = G.dirty2vis(uvw, freq, dirty, wgt, px, py, wg, mask, **kw)
    rv = R.wgridder(True, uvw, freq, dirty, wgt, mask, nx, ny, px, py,
                    1e-10, wg, **kw)
    assert O.l2error(v, rv) <= 1e-10
