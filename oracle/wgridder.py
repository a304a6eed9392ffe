"""CPU restatement of ducc0's w-gridder as direct sums (test infrastructure
only: used by tests/ as the checker, never by the library).

vis2dirty / ms2dirty (src/ducc0/wgridder/wgridder.h:1626-1700, x2dirty at
:1429-1459) compute, for the pixels (i, j) of an npix_x x npix_y image at
    x = center_x + (i - npix_x/2) pixsize_x,  y = center_y + (j - npix_y/2) pixsize_y
(with u -> -u, center_x -> -center_x under flip_u, likewise v and w),
    dirty(i, j) = sum_{row, chan} Re(vis wgt exp(2 pi i (x u + y v - w (n - 1)))) / n
where (u, v, w) = uvw[row] * freq[chan] / c, n = sqrt(1 - x^2 - y^2), and
without w-gridding n - 1 := 0 and no division; pixels beyond the horizon
(x^2 + y^2 > 1) get n - 1 = -sqrt(x^2 + y^2 - 1) - 1 and, when dividing by
n, zero (wgridder.h:1361-1371).  dirty2vis / dirty2ms is the adjoint:
    vis(row, chan) = wgt sum_{i, j} dirty(i, j) / n exp(-2 pi i (x u + y v - w (n - 1))).
Masked (mask == 0) visibilities contribute nothing and come back zero.
This is the sum the reference's own tests check against
(python/test/test_wgridder.py:34-91), written out for the test sizes.
Pinned in tests/test_wgridder_oracle.py: without w-gridding and without
shifts it is a 2-D type-1 NUDFT (oracle.core.nudft_type1), the pair is
adjoint, and ducc0 itself (oracle/_ref) agrees."""
import numpy as np

SPEED_OF_LIGHT = 299792458.


def _pixels(nx, ny, px, py, cx, cy):
    x = cx + (np.arange(nx) - nx // 2) * px
    y = cy + (np.arange(ny) - ny // 2) * py
    X, Y = np.meshgrid(x, y, indexing="ij")
    return X, Y


def _nm1_n(X, Y, apply_w):
    if not apply_w:
        return np.zeros_like(X), np.ones_like(X)
    r2 = X * X + Y * Y
    inside = r2 <= 1.
    nm1 = np.where(inside, -r2 / (np.sqrt(np.abs(1. - r2)) + 1.),
                   -np.sqrt(np.abs(r2 - 1.)) - 1.)
    return nm1, nm1 + 1.


def _eff(uvw, freq, flip_u, flip_v, flip_w):
    f = np.array([-1. if flip_u else 1., -1. if flip_v else 1.,
                  -1. if flip_w else 1.])
    return (uvw * f)[:, None, :] * (np.asarray(freq)[None, :, None] /
                                    SPEED_OF_LIGHT)


def vis2dirty(uvw, freq, vis, wgt, npix_x, npix_y, pixsize_x, pixsize_y,
              do_wgridding, mask=None, flip_u=False, flip_v=False,
              flip_w=False, divide_by_n=True, center_x=0., center_y=0.):
    cx = -center_x if flip_u else center_x
    cy = -center_y if flip_v else center_y
    X, Y = _pixels(npix_x, npix_y, pixsize_x, pixsize_y, cx, cy)
    nm1, n = _nm1_n(X, Y, do_wgridding)
    uvwe = _eff(np.asarray(uvw, np.float64), freq, flip_u, flip_v, flip_w)
    v = np.asarray(vis, np.complex128)
    if wgt is not None:
        v = v * wgt
    if mask is not None:
        v = np.where(np.asarray(mask) != 0, v, 0)
    res = np.zeros((npix_x, npix_y))
    for r in range(v.shape[0]):
        for c in range(v.shape[1]):
            if v[r, c] == 0:
                continue
            u, vv, w = uvwe[r, c]
            ph = 2j * np.pi * (X * u + Y * vv - w * nm1)
            res += (v[r, c] * np.exp(ph)).real
    if do_wgridding and divide_by_n:
        res = np.where(n > 0, res / np.where(n > 0, n, 1.), 0.)
    return res


def dirty2vis(uvw, freq, dirty, wgt, pixsize_x, pixsize_y, do_wgridding,
              mask=None, flip_u=False, flip_v=False, flip_w=False,
              divide_by_n=True, center_x=0., center_y=0.):
    npix_x, npix_y = dirty.shape
    cx = -center_x if flip_u else center_x
    cy = -center_y if flip_v else center_y
    X, Y = _pixels(npix_x, npix_y, pixsize_x, pixsize_y, cx, cy)
    nm1, n = _nm1_n(X, Y, do_wgridding)
    d = np.asarray(dirty, np.float64)
    if do_wgridding and divide_by_n:
        d = np.where(n > 0, d / np.where(n > 0, n, 1.), 0.)
    uvwe = _eff(np.asarray(uvw, np.float64), freq, flip_u, flip_v, flip_w)
    nrow, nchan = uvwe.shape[:2]
    res = np.zeros((nrow, nchan), np.complex128)
    for r in range(nrow):
        for c in range(nchan):
            if mask is not None and mask[r, c] == 0:
                continue
            u, vv, w = uvwe[r, c]
            ph = -2j * np.pi * (X * u + Y * vv - w * nm1)
            res[r, c] = np.sum(d * np.exp(ph))
    if wgt is not None:
        res = res * wgt
    return res
