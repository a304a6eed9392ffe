"""ducc0.nufft-compatible surface backed by the MI355X kernels.

Same names, keyword-only arguments, defaults, dtype acceptance and error
messages as ducc0.nufft (python/nufft_pymod.cc:445-471), plus a few keyword
arguments ducc0 does not have (all with defaults that keep ducc0's
behaviour apart from the compute precision, see `u2nu`):

  precision  "auto" | "fp32" | "fp64".  "auto" computes in the precision
             of the values (as ducc0 does), raised to FP64 for complex64
             values when epsilon < 1e-6 or the single-precision kernel table
             cannot reach epsilon; coordinates are always folded in FP64 and
             correction factors are FP64.
  chooser    "analytic" (ducc0's cost model) | "learned" (the trained
             parameter network, with the analytic choice as fallback).
  deterministic
             True: type-1 results identical bit for bit from run to run
             (spreading without floating-point atomics; slower).  Type 2
             is always deterministic.
  point_records
             True: unplanned calls sort the points' records instead of a
             permutation (same results; slower on MI355X).
  fp64_accumulation
             True: type-1 spreading of the FP32 path accumulates each tile
             in FP64 (two LDS float atomics per tap) instead of packed
             fixed point (one 64-bit integer atomic per tap, used at
             eps >= 1e-5).
  check_offsets
             True: spreading / interpolation count points whose fold lands
             outside their tile and the call fails if any did (a debugging
             aid: one counter and a wait per call).
  info       a dict that receives the parameters actually used.

Arrays may be numpy arrays (host memory: copied to the GPU and back) or
torch tensors on the GPU (used in place, on torch's current stream).
`nthreads` is accepted and ignored.
"""
import ctypes as ct

import numpy as np

from . import _lib

__all__ = ["u2nu", "nu2u", "plan", "bestEpsilon"]

_TWO_PI = 2 * np.pi


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _periodicity(p, ndim):
    """python/nufft_pymod.cc:36-50"""
    try:
        v = float(p)
        return np.full(ndim, v, dtype=np.float64)
    except TypeError:
        arr = np.asarray(p, dtype=np.float64).ravel()
        if arr.size != ndim:
            raise RuntimeError("bad size of periodicity argument")
        return np.ascontiguousarray(arr)


def _stream_for(*arrays):
    for a in arrays:
        if _is_torch(a) and a.is_cuda:
            import torch
            return torch.cuda.current_stream(a.device).cuda_stream
    return None


def _check_dtypes(vals, coord):
    """unplanned calls accept coords {f64,f32} x values {c128,c64}
    (python/nufft_pymod.cc:70-94)"""
    if _is_torch(vals) or _is_torch(coord):
        import torch
        okc = coord.dtype in (torch.float32, torch.float64)
        okv = vals.dtype in (torch.complex64, torch.complex128)
    else:
        okc = np.asarray(coord).dtype in (np.float32, np.float64)
        okv = np.asarray(vals).dtype in (np.complex64, np.complex128)
    if not (okc and okv):
        raise RuntimeError("not yet supported")


def _opts(precision, chooser, log2tile, stream, kernel_idx=-1, sigma=0.0,
          spot_check=True, deterministic=False, point_records=False,
          check_offsets=False, fp64_accumulation=False):
    return _lib.options(precision=precision, chooser=chooser,
                        log2tile=log2tile, stream=stream,
                        kernel_idx=kernel_idx, sigma=sigma,
                        spot_check=spot_check, deterministic=deterministic,
                        point_records=point_records,
                        check_offsets=check_offsets,
                        fp64_accumulation=fp64_accumulation)


def u2nu(*, grid, coord, forward, epsilon, nthreads=1, out=None,
         verbosity=0, sigma_min=1.2, sigma_max=2.51,
         periodicity=_TWO_PI, fft_order=False, precision="auto",
         chooser="analytic", log2tile=0, info=None, kernel_idx=-1,
         spot_check=True, deterministic=False, point_records=False,
         check_offsets=False):
    """Type 2 NUFFT: uniform grid -> values at nonuniform points.

    out_j = sum_k grid_k exp(isign i 2pi/P k.x_j), isign = -1 if forward,
    modes k from -(n//2) to n-1-(n//2) (or FFT order).  Returns `out`
    (the same object when given)."""
    _lib.require_gpu()
    _check_dtypes(grid, coord)
    nd = coord.shape[1] if len(coord.shape) == 2 else 0
    per = _periodicity(periodicity, max(nd, 1))
    if out is None:
        if _is_torch(grid):
            import torch
            out = torch.empty(coord.shape[0], dtype=grid.dtype,
                              device=grid.device)
        else:
            out = np.empty(coord.shape[0], dtype=np.asarray(grid).dtype)
    g, kg = _lib.describe(grid)
    c, kc = _lib.describe(coord)
    o, ko = _lib.describe(out)
    opt = _opts(precision, chooser, log2tile, _stream_for(grid, coord, out),
                kernel_idx=kernel_idx, spot_check=spot_check,
                deterministic=deterministic, point_records=point_records,
                check_offsets=check_offsets)
    inf = _lib.DmxPlanInfo()
    _lib.check(_lib.lib().dmx_u2nu(
        ct.byref(g), ct.byref(c), int(bool(forward)), float(epsilon),
        int(nthreads), ct.byref(o), int(verbosity), float(sigma_min),
        float(sigma_max), per.ctypes.data, int(bool(fft_order)),
        ct.byref(opt), ct.byref(inf)))
    if info is not None:
        info.update(inf.as_dict())
    if not _is_torch(out) and ko is not out:
        out[...] = ko
    return out


def nu2u(*, points, coord, forward, epsilon, nthreads=1, out=None,
         verbosity=0, sigma_min=1.2, sigma_max=2.51,
         periodicity=_TWO_PI, fft_order=False, precision="auto",
         chooser="analytic", log2tile=0, info=None, kernel_idx=-1,
         spot_check=True, deterministic=False, point_records=False,
         check_offsets=False, fp64_accumulation=False):
    """Type 1 NUFFT: values at nonuniform points -> uniform grid.

    out_k = sum_j points_j exp(isign i 2pi/P k.x_j).  `out` is required: its
    shape defines the uniform grid (as in ducc0)."""
    if out is None:
        raise TypeError("nu2u(): 'out' is required, its shape defines the "
                        "uniform grid")
    _lib.require_gpu()
    _check_dtypes(points, coord)
    nd = coord.shape[1] if len(coord.shape) == 2 else 0
    per = _periodicity(periodicity, max(nd, 1))
    p, kp = _lib.describe(points)
    c, kc = _lib.describe(coord)
    o, ko = _lib.describe(out)
    opt = _opts(precision, chooser, log2tile, _stream_for(points, coord, out),
                kernel_idx=kernel_idx, spot_check=spot_check,
                deterministic=deterministic, point_records=point_records,
                check_offsets=check_offsets,
                fp64_accumulation=fp64_accumulation)
    inf = _lib.DmxPlanInfo()
    _lib.check(_lib.lib().dmx_nu2u(
        ct.byref(p), ct.byref(c), int(bool(forward)), float(epsilon),
        int(nthreads), ct.byref(o), int(verbosity), float(sigma_min),
        float(sigma_max), per.ctypes.data, int(bool(fft_order)),
        ct.byref(opt), ct.byref(inf)))
    if info is not None:
        info.update(inf.as_dict())
    if not _is_torch(out) and ko is not out:
        out[...] = ko
    return out


def u2nu_coord_grad(*, grid, coord, forward, epsilon, out=None,
                    sigma_min=1.2, sigma_max=2.51, periodicity=_TWO_PI,
                    fft_order=False, precision="auto", chooser="analytic",
                    log2tile=0):
    """d u2nu(grid)_j / d coord[j, d] for every point and axis, as an
    (M, ndim) complex array: the interpolation of the same oversampled grid
    with the derivative of the spreading kernel (no ducc0 counterpart)."""
    _lib.require_gpu()
    _check_dtypes(grid, coord)
    M, nd = coord.shape
    per = _periodicity(periodicity, nd)
    if out is None:
        if _is_torch(grid):
            import torch
            out = torch.empty((M, nd), dtype=grid.dtype, device=grid.device)
        else:
            out = np.empty((M, nd), dtype=np.asarray(grid).dtype)
    g, kg = _lib.describe(grid)
    c, kc = _lib.describe(coord)
    o, ko = _lib.describe(out)
    opt = _opts(precision, chooser, log2tile, _stream_for(grid, coord, out))
    _lib.check(_lib.lib().dmx_u2nu_coord_grad(
        ct.byref(g), ct.byref(c), int(bool(forward)), float(epsilon),
        ct.byref(o), float(sigma_min), float(sigma_max), per.ctypes.data,
        int(bool(fft_order)), ct.byref(opt)))
    if not _is_torch(out) and ko is not out:
        out[...] = ko
    return out


def bestEpsilon(*, ndim, singleprec, sigma_min=1.1, sigma_max=2.6):
    """Smallest achievable epsilon (gridding_kernel.cc:1562-1574)."""
    r = _lib.lib().dmx_best_epsilon(int(ndim), int(bool(singleprec)),
                                    float(sigma_min), float(sigma_max))
    if r < 0:
        raise RuntimeError(_lib.lib().dmx_last_error().decode())
    return r


class plan:
    """Pre-planned NUFFT: the points are sorted once (python/nufft_pymod.cc
    :139-261).  Values must have the precision of the coordinates."""

    def __init__(self, *, nu2u, coord, grid_shape, epsilon, nthreads=0,
                 sigma_min=1.1, sigma_max=2.6, periodicity=_TWO_PI,
                 fft_order=False, precision="auto", chooser="analytic",
                 log2tile=0, deterministic=False, check_offsets=False):
        _lib.require_gpu()
        self._h = None
        if _is_torch(coord):
            import torch
            if coord.dtype not in (torch.float32, torch.float64):
                raise RuntimeError("not yet supported")
            self._single = coord.dtype == torch.float32
        else:
            if np.asarray(coord).dtype not in (np.float32, np.float64):
                raise RuntimeError("not yet supported")
            self._single = np.asarray(coord).dtype == np.float32
        self._shape = tuple(int(s) for s in grid_shape)
        nd = len(self._shape)
        per = _periodicity(periodicity, nd)
        c, kc = _lib.describe(coord)
        shp = np.asarray(self._shape, dtype=np.int64)
        self._stream = _stream_for(coord)
        opt = _opts(precision, chooser, log2tile, self._stream,
                    deterministic=deterministic, check_offsets=check_offsets)
        h = ct.c_void_p()
        _lib.check(_lib.lib().dmx_plan_create(
            ct.byref(h), int(bool(nu2u)), ct.byref(c), shp.ctypes.data, nd,
            float(epsilon), int(nthreads), float(sigma_min), float(sigma_max),
            per.ctypes.data, int(bool(fft_order)), ct.byref(opt)))
        self._h = h
        self._npoints = coord.shape[0]

    def _vdtype_ok(self, x):
        if _is_torch(x):
            import torch
            want = torch.complex64 if self._single else torch.complex128
        else:
            want = np.complex64 if self._single else np.complex128
            x = np.asarray(x)
        if x.dtype != want:
            raise RuntimeError("not yet supported")

    @staticmethod
    def _call_opts(*arrays):
        """Options of one planned call: torch's current stream when a
        device tensor takes part (the call is then ordered on it like a
        torch operation), else None (the plan's own stream, waiting for the
        results)."""
        s = _stream_for(*arrays)
        if s is None:
            return None
        return ct.byref(_lib.options(stream=s))

    def nu2u(self, *, forward, verbosity=0, points, out=None):
        self._vdtype_ok(points)
        if out is None:
            if _is_torch(points):
                import torch
                out = torch.empty(self._shape, dtype=points.dtype,
                                  device=points.device)
            else:
                out = np.empty(self._shape, dtype=np.asarray(points).dtype)
        p, kp = _lib.describe(points)
        o, ko = _lib.describe(out)
        _lib.check(_lib.lib().dmx_plan_nu2u(self._h, int(bool(forward)),
                                            int(verbosity), ct.byref(p),
                                            ct.byref(o),
                                            self._call_opts(points, out)))
        if not _is_torch(out) and ko is not out:
            out[...] = ko
        return out

    def u2nu(self, *, forward, verbosity=0, grid, out=None):
        self._vdtype_ok(grid)
        if out is None:
            if _is_torch(grid):
                import torch
                out = torch.empty(self._npoints, dtype=grid.dtype,
                                  device=grid.device)
            else:
                out = np.empty(self._npoints, dtype=np.asarray(grid).dtype)
        g, kg = _lib.describe(grid)
        o, ko = _lib.describe(out)
        _lib.check(_lib.lib().dmx_plan_u2nu(self._h, int(bool(forward)),
                                            int(verbosity), ct.byref(g),
                                            ct.byref(o),
                                            self._call_opts(grid, out)))
        if not _is_torch(out) and ko is not out:
            out[...] = ko
        return out

    def info(self):
        inf = _lib.DmxPlanInfo()
        _lib.check(_lib.lib().dmx_plan_get_info(self._h, ct.byref(inf)))
        return inf.as_dict()

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                _lib.lib().dmx_plan_destroy(self._h)
            except Exception:
                pass
            self._h = None


def nudft_direct(type_, coord, values, shape, forward, sel,
                 periodicity=_TWO_PI, fft_order=False):
    """Direct FP64 NUDFT on the GPU for selected modes (type 1) or points
    (type 2): the spot-check of the parameter chooser."""
    _lib.require_gpu()
    coord = np.ascontiguousarray(coord, np.float64)
    values = np.ascontiguousarray(values, np.complex128)
    shape = tuple(int(s) for s in shape)
    sel = np.ascontiguousarray(sel, np.int64)
    out = np.zeros(len(sel), np.complex128)
    per = _periodicity(periodicity, len(shape))
    c, _ = _lib.describe(coord)
    v, _ = _lib.describe(values)
    o, _ = _lib.describe(out)
    shp = np.asarray(shape, np.int64)
    _lib.check(_lib.lib().dmx_nudft(int(type_), ct.byref(c), ct.byref(v),
                                    shp.ctypes.data, len(shape),
                                    int(bool(forward)), per.ctypes.data,
                                    int(bool(fft_order)), sel.ctypes.data,
                                    len(sel), ct.byref(o)))
    return out


def sort_records(coord, nover, W, log2tile, periodicity=_TWO_PI):
    """The record sort of unplanned transforms on the GPU: per sorted
    position the record word (tile << ndim*log2tile | tile-local offsets),
    the point's input index and its kernel arguments W-1-2*frac (float64),
    for bit-exact parity checks."""
    _lib.require_gpu()
    coord = np.ascontiguousarray(coord)
    M, nd = coord.shape
    words = np.zeros(M, np.uint32)
    idx = np.zeros(M, np.uint32)
    xarg = np.zeros((M, nd), np.float64)
    per = _periodicity(periodicity, nd)
    c, _ = _lib.describe(coord)
    w, _ = _lib.describe(words)
    i, _ = _lib.describe(idx)
    x, _ = _lib.describe(xarg)
    nov = np.asarray(nover, np.int64)
    _lib.check(_lib.lib().dmx_sort_records(ct.byref(c), nov.ctypes.data, nd,
                                           int(W), int(log2tile),
                                           per.ctypes.data, ct.byref(w),
                                           ct.byref(i), ct.byref(x)))
    return words, idx, xarg


def sort_points(coord, nover, W, log2tile, periodicity=_TWO_PI):
    """Tile keys (input order) and the stable sort permutation computed on
    the GPU (for bit-exact parity checks of the index build)."""
    _lib.require_gpu()
    coord = np.ascontiguousarray(coord)
    M, nd = coord.shape
    keys = np.zeros(M, np.uint32)
    idx = np.zeros(M, np.uint32)
    per = _periodicity(periodicity, nd)
    c, _ = _lib.describe(coord)
    k, _ = _lib.describe(keys)
    i, _ = _lib.describe(idx)
    nov = np.asarray(nover, np.int64)
    _lib.check(_lib.lib().dmx_sort_points(ct.byref(c), nov.ctypes.data, nd,
                                          int(W), int(log2tile),
                                          per.ctypes.data, ct.byref(k),
                                          ct.byref(i)))
    return keys, idx
