"""ducc0.wgridder on the GPU: the Python surface of vis2dirty / dirty2vis
(keyword-only) and the older ms2dirty / dirty2ms (positional), with ducc0's
names and defaults (python/wgridder_pymod.cc:483-560, 612-622).

Arrays may be numpy (copied to and from the device) or torch tensors on the
GPU (used in place, ordered on torch's current stream).  The precision of a
call is that of `vis` (vis2dirty) or `dirty` (dirty2vis): complex64 /
float32 or complex128 / float64.  `gpu`, `nthreads` and
`double_precision_accumulation` are accepted for compatibility (the work
always runs on the GPU; each tile's sums are accumulated in FP64 in LDS,
and the w-plane sums of the image in FP64)."""
import ctypes as ct

import numpy as np

from . import _lib
from ._lib import _is_torch


def _opts(*arrays):
    if any(_is_torch(a) and a.is_cuda for a in arrays):
        import torch
        return _lib.options(stream=torch.cuda.current_stream().cuda_stream)
    return _lib.options()


def _keep(*xs):
    """descriptors whose arrays must stay referenced for the call"""
    out = []
    for x in xs:
        if x is None:
            out.append((None, None))
        else:
            d, ref = _lib.describe(x)
            out.append((d, ref))
    return out


def _real_dtype_of(vis):
    if _is_torch(vis):
        import torch
        return torch.float32 if vis.dtype == torch.complex64 else torch.float64
    return np.float32 if np.asarray(vis).dtype == np.complex64 else np.float64


def _cplx_dtype_of(dirty):
    if _is_torch(dirty):
        import torch
        return torch.complex64 if dirty.dtype == torch.float32 \
            else torch.complex128
    return np.complex64 if np.asarray(dirty).dtype == np.float32 \
        else np.complex128


def _mask_u8(mask):
    if mask is None or _is_torch(mask):
        return mask
    return np.ascontiguousarray(mask, dtype=np.uint8)


def vis2dirty(*, uvw, freq, vis, wgt=None, npix_x=0, npix_y=0, pixsize_x,
              pixsize_y, epsilon, do_wgridding=False, nthreads=1, verbosity=0,
              mask=None, flip_u=False, flip_v=False, flip_w=False,
              divide_by_n=True, dirty=None, sigma_min=1.1, sigma_max=2.6,
              center_x=0., center_y=0., allow_nshift=True, gpu=False,
              double_precision_accumulation=False):
    """Dirty image from visibilities (wgridder.h:1700-1714).  Returns
    `dirty` (the same object when given), shape (npix_x, npix_y)."""
    _lib.require_gpu()
    mask = _mask_u8(mask)
    if dirty is None:
        if npix_x <= 0 or npix_y <= 0:
            raise RuntimeError("need npix_x and npix_y or dirty")
        if _is_torch(vis):
            import torch
            dirty = torch.empty((npix_x, npix_y), dtype=_real_dtype_of(vis),
                                device=vis.device)
        else:
            dirty = np.empty((npix_x, npix_y), dtype=_real_dtype_of(vis))
    (u, _u), (f, _f), (v, _v), (w, _w), (m, _m), (d, _d) = _keep(
        uvw, freq, vis, wgt, mask, dirty)
    opt = _opts(uvw, vis, dirty)
    _lib.check(_lib.lib().dmx_vis2dirty(
        ct.byref(u), ct.byref(f), ct.byref(v),
        None if w is None else ct.byref(w), None if m is None else ct.byref(m),
        float(pixsize_x), float(pixsize_y), float(epsilon),
        int(bool(do_wgridding)), int(bool(flip_u)), int(bool(flip_v)),
        int(bool(flip_w)), int(bool(divide_by_n)), float(sigma_min),
        float(sigma_max), float(center_x), float(center_y),
        int(bool(allow_nshift)), int(nthreads), int(verbosity), ct.byref(d),
        ct.byref(opt)))
    return dirty


def dirty2vis(*, uvw, freq, dirty, wgt=None, pixsize_x, pixsize_y, epsilon,
              do_wgridding=False, nthreads=1, verbosity=0, mask=None,
              flip_u=False, flip_v=False, flip_w=False, divide_by_n=True,
              vis=None, sigma_min=1.1, sigma_max=2.6, center_x=0.,
              center_y=0., allow_nshift=True, gpu=False):
    """Visibilities from a dirty image (wgridder.h:1716-1735), the adjoint
    of vis2dirty.  Returns `vis` (the same object when given), shape
    (nrow, nchan)."""
    _lib.require_gpu()
    mask = _mask_u8(mask)
    if vis is None:
        shape = (uvw.shape[0], freq.shape[0])
        if _is_torch(dirty):
            import torch
            vis = torch.empty(shape, dtype=_cplx_dtype_of(dirty),
                              device=dirty.device)
        else:
            vis = np.empty(shape, dtype=_cplx_dtype_of(dirty))
    (u, _u), (f, _f), (d, _d), (w, _w), (m, _m), (v, _v) = _keep(
        uvw, freq, dirty, wgt, mask, vis)
    opt = _opts(uvw, vis, dirty)
    _lib.check(_lib.lib().dmx_dirty2vis(
        ct.byref(u), ct.byref(f), ct.byref(d),
        None if w is None else ct.byref(w), None if m is None else ct.byref(m),
        float(pixsize_x), float(pixsize_y), float(epsilon),
        int(bool(do_wgridding)), int(bool(flip_u)), int(bool(flip_v)),
        int(bool(flip_w)), int(bool(divide_by_n)), float(sigma_min),
        float(sigma_max), float(center_x), float(center_y),
        int(bool(allow_nshift)), int(nthreads), int(verbosity), ct.byref(v),
        ct.byref(opt)))
    return vis


def ms2dirty(uvw, freq, ms, wgt=None, npix_x=0, npix_y=0, pixsize_x=None,
             pixsize_y=None, nu=0, nv=0, epsilon=None, do_wstacking=False,
             nthreads=1, verbosity=0, mask=None,
             double_precision_accumulation=False, gpu=False):
    """ms2dirty (wgridder_pymod.cc:483-493): vis2dirty with nu / nv ignored
    and the default options"""
    return vis2dirty(uvw=uvw, freq=freq, vis=ms, wgt=wgt, npix_x=npix_x,
                     npix_y=npix_y, pixsize_x=pixsize_x, pixsize_y=pixsize_y,
                     epsilon=epsilon, do_wgridding=do_wstacking,
                     nthreads=nthreads, verbosity=verbosity, mask=mask)


def dirty2ms(uvw, freq, dirty, wgt=None, pixsize_x=None, pixsize_y=None, nu=0,
             nv=0, epsilon=None, do_wstacking=False, nthreads=1, verbosity=0,
             mask=None, gpu=False):
    """dirty2ms (wgridder_pymod.cc:544-550): dirty2vis with nu / nv ignored
    and the default options"""
    return dirty2vis(uvw=uvw, freq=freq, dirty=dirty, wgt=wgt,
                     pixsize_x=pixsize_x, pixsize_y=pixsize_y,
                     epsilon=epsilon, do_wgridding=do_wstacking,
                     nthreads=nthreads, verbosity=verbosity, mask=mask)
