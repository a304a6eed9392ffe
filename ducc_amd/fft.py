"""ducc0.fft.c2c-compatible complex FFT on the MI355X Stockham kernels.

Same name, arguments, defaults, dtype rules and error texts as ducc0.fft.c2c
(python/fft_pymod.cc:73-94, 104-172, 975-976).  Arrays may be numpy arrays
(host memory: copied to the GPU and back) or torch tensors on the GPU (used
in place, on torch's current stream, returning without waiting as torch
operations do).  `nthreads` is accepted and ignored.
"""
import ctypes as ct
import math

import numpy as np

from . import _lib

__all__ = ["c2c", "good_size"]


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def good_size(n, real=False):
    """Smallest length >= n that is 2,3,5,7,11-smooth (2,3,5-smooth if
    `real`): ducc0.fft.good_size (src/ducc0/fft/fft.h:93-119, the complex
    and real variants; argument checks of python/fft_pymod.cc:427-449)."""
    n = int(n)
    if n < 0:
        raise ValueError("Target length must be positive")
    if n <= 6:
        return n
    primes = (2, 3, 5) if real else (2, 3, 5, 7, 11)
    while True:
        m = n
        for p in primes:
            while m % p == 0:
                m //= p
        if m == 1:
            return n
        n += 1


def _makeaxes(ndim, axes):
    """python/fft_pymod.cc:73-94"""
    if axes is None:
        return list(range(ndim))
    tmp = [int(a) for a in (axes if np.iterable(axes) else [axes])]
    if len(tmp) > ndim or len(tmp) == 0:
        raise RuntimeError("bad axes argument")
    res = []
    for a in tmp:
        if a < 0:
            a += ndim
        if a >= ndim or a < 0:
            raise ValueError("axes exceeds dimensionality of output")
        res.append(a)
    return res


def _norm_fct(inorm, shape, axes):
    """python/fft_pymod.cc:104-120"""
    if inorm == 0:
        return 1.0
    N = 1
    for a in axes:
        N *= int(shape[a])
    if inorm == 2:
        return 1.0 / N
    if inorm == 1:
        return 1.0 / math.sqrt(N)
    raise ValueError("invalid value for inorm (must be 0, 1, or 2)")


def c2c(a, axes=None, forward=True, inorm=0, out=None, nthreads=1):
    """Complex FFT of `a` along `axes` (all axes if None); see
    ducc0.fft.c2c (python/fft_pymod.cc:493-525).  Real input is transformed
    as complex input with zero imaginary part; the result is complex of the
    input's precision."""
    _lib.require_gpu()
    if _is_torch(a):
        import torch
        cmap = {torch.complex64: torch.complex64,
                torch.complex128: torch.complex128,
                torch.float32: torch.complex64,
                torch.float64: torch.complex128}
        if a.dtype not in cmap:
            raise RuntimeError("unsupported data type")
        odt = cmap[a.dtype]
        shape = tuple(a.shape)
        if out is None:
            out = torch.empty(shape, dtype=odt, device=a.device)
        elif not _is_torch(out) or out.dtype != odt:
            raise RuntimeError("incorrect data type")
        stream = (torch.cuda.current_stream(a.device).cuda_stream
                  if a.is_cuda and out.is_cuda else None)
    else:
        a = np.asarray(a)
        cmap = {np.dtype(np.complex64): np.complex64,
                np.dtype(np.complex128): np.complex128,
                np.dtype(np.float32): np.complex64,
                np.dtype(np.float64): np.complex128}
        if a.dtype not in cmap:
            raise RuntimeError("unsupported data type")
        odt = cmap[a.dtype]
        shape = a.shape
        if out is None:
            out = np.empty(shape, dtype=odt)
        elif _is_torch(out) or np.asarray(out).dtype != np.dtype(odt):
            raise RuntimeError("incorrect data type")
        stream = None
    if len(tuple(out.shape)) != len(shape):
        raise RuntimeError("dimension mismatch")
    if tuple(out.shape) != tuple(shape):
        raise RuntimeError("dimension mismatch")
    ax = _makeaxes(len(shape), axes)
    fct = _norm_fct(inorm, shape, ax)
    if len(shape) == 0:  # a 0-d array: a length-1 transform
        c2c(a.reshape(1), axes=None, forward=forward, inorm=inorm,
            out=out.reshape(1))
        return out
    # host arrays are staged through device copies (the input before the
    # output is touched), so out may be a itself
    da, ka = _lib.describe(a)
    do, ko = _lib.describe(out)
    axv = np.asarray(ax, dtype=np.int64)
    opt = _lib.options(stream=stream)
    _lib.check(_lib.lib().dmx_c2c(ct.byref(da), ct.byref(do),
                                  axv.ctypes.data, len(ax), int(bool(forward)),
                                  float(fct), int(nthreads), ct.byref(opt)))
    if not _is_torch(out) and ko is not out:
        out[...] = ko
    return out
