"""ctypes wrapper over oracle/_ref/libducc_ref.so: the reference ducc0 NUFFT
compiled from its own sources (oracle/Makefile target `ref`).

TEST INFRASTRUCTURE ONLY.  `available()` is False where the library was not
built (it needs the ducc0 sources at build time); callers skip then.
"""
import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_ref", "libducc_ref.so")
_lib = None


def available():
    return os.path.exists(LIB)


def lib():
    global _lib
    if _lib is None:
        L = ct.CDLL(LIB)
        P, I, D, S = ct.c_void_p, ct.c_int, ct.c_double, ct.c_size_t
        L.ref_last_error.restype = ct.c_char_p
        L.ref_vlen.restype = I
        L.ref_vlen.argtypes = [I]
        L.ref_best_epsilon.restype = D
        L.ref_best_epsilon.argtypes = [I, I, D, D]
        L.ref_find_params.restype = I
        L.ref_find_params.argtypes = [D, D, D, I, P, S, I, I, S, P]
        L.ref_kernel_coeffs.argtypes = [I, P, P]
        L.ref_corfac.argtypes = [I, S, D, P]
        L.ref_sort_index.argtypes = [I, S, P, P, P, D, D, D, I, I, S, P, P, P]
        L.ref_nu2u.argtypes = [I, S, P, I, P, I, I, D, S, P, P, D, D, P, I]
        L.ref_u2nu.argtypes = L.ref_nu2u.argtypes
        L.ref_c2c.argtypes = [I, I, P, P, P, P, I, I, D, S]
        L.ref_plan_create.argtypes = [ct.POINTER(P), I, I, S, P, I, P, D, S,
                                      D, D, P, I]
        L.ref_plan_nu2u.argtypes = [P, I, P, P]
        L.ref_plan_u2nu.argtypes = [P, I, P, P]
        L.ref_plan_destroy.argtypes = [P]
        L.ref_plan_destroy.restype = None
        L.ref_sht_general.argtypes = [I, I, P, S, P, S, P, S, ct.c_ssize_t, P,
                                      S, D, D, D, S, S, I]
        L.ref_wgridder.argtypes = [I, I, P, S, P, S, P, P, P, P, S, S, D, D,
                                   D, I, I, I, I, I, D, D, D, D, S]
        L.ref_sht_pseudo_analysis.argtypes = [I, P, S, P, S, P, S,
                                              ct.c_ssize_t, P, S, D, D, S, S,
                                              D, S, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ct.c_void_p)


def _check(rc):
    if rc:
        raise RuntimeError(lib().ref_last_error().decode())


def vlen(single):
    return lib().ref_vlen(int(single))


def best_epsilon(ndim, single, sigma_min=1.1, sigma_max=2.6):
    r = lib().ref_best_epsilon(ndim, int(single), sigma_min, sigma_max)
    if r < 0:
        raise RuntimeError(lib().ref_last_error().decode())
    return r


def find_params(eps, shape, npoints, gridding, single, nthreads,
                sigma_min=1.2, sigma_max=2.51):
    dims = np.asarray(shape, np.uintp)
    nover = np.zeros(3, np.uintp)
    k = lib().ref_find_params(eps, sigma_min, sigma_max, len(shape), _p(dims),
                              npoints, int(gridding), int(single), nthreads,
                              _p(nover))
    if k < 0:
        raise RuntimeError(lib().ref_last_error().decode())
    return k, tuple(int(x) for x in nover[:len(shape)])


def kernel_coeffs(kidx, W):
    c = np.zeros(64 * 20)
    D = np.zeros(1, np.int32)
    _check(lib().ref_kernel_coeffs(kidx, _p(c), _p(D)))
    d = int(D[0])
    return c[:(d + 1) * W].reshape(d + 1, W)


def corfac(kidx, n, dx):
    out = np.zeros(n)
    _check(lib().ref_corfac(kidx, n, dx, _p(out)))
    return out


def sort_index(coords, shape, eps, gridding, single, nthreads=1,
               period=2 * np.pi, sigma_min=1.2, sigma_max=2.51):
    """coord_idx of a reference plan -> (idx, nover, supp)"""
    c = np.ascontiguousarray(coords, np.float64)
    M, d = c.shape
    idx = np.zeros(M, np.uint32)
    nover = np.zeros(3, np.uintp)
    supp = np.zeros(1, np.int32)
    per = np.ascontiguousarray(np.broadcast_to(np.asarray(period, float),
                                               (d,)))
    _check(lib().ref_sort_index(d, M, _p(c), _p(np.asarray(shape, np.uintp)),
                                _p(per), eps, sigma_min, sigma_max,
                                int(gridding), int(single), nthreads, _p(idx),
                                _p(nover), _p(supp)))
    return idx, tuple(int(x) for x in nover[:d]), int(supp[0])


def _prep(coords, vals, period):
    coord_single = coords.dtype == np.float32
    c = np.ascontiguousarray(coords,
                             np.float32 if coord_single else np.float64)
    single = vals.dtype == np.complex64
    v = np.ascontiguousarray(vals, np.complex64 if single else np.complex128)
    d = c.shape[1]
    per = np.ascontiguousarray(np.broadcast_to(np.asarray(period, float),
                                               (d,)))
    return c, int(coord_single), v, int(single), per


def nu2u(points, coord, forward, epsilon, out_shape, nthreads=1,
         sigma_min=1.2, sigma_max=2.51, periodicity=2 * np.pi,
         fft_order=False):
    """ducc0.nufft.nu2u semantics (python/nufft_pymod.cc:96-137)"""
    c, cs, v, s, per = _prep(coord, points, periodicity)
    out = np.zeros(out_shape, v.dtype)
    _check(lib().ref_nu2u(c.shape[1], c.shape[0], _p(c), cs, _p(v), s,
                          int(forward), epsilon, nthreads,
                          _p(np.asarray(out_shape, np.uintp)), _p(out),
                          sigma_min, sigma_max, _p(per), int(fft_order)))
    return out


def u2nu(grid, coord, forward, epsilon, nthreads=1, sigma_min=1.2,
         sigma_max=2.51, periodicity=2 * np.pi, fft_order=False):
    """ducc0.nufft.u2nu semantics (python/nufft_pymod.cc:52-94)"""
    c, cs, v, s, per = _prep(coord, grid, periodicity)
    out = np.zeros(c.shape[0], v.dtype)
    _check(lib().ref_u2nu(c.shape[1], c.shape[0], _p(c), cs, _p(v), s,
                          int(forward), epsilon, nthreads,
                          _p(np.asarray(v.shape, np.uintp)), _p(out),
                          sigma_min, sigma_max, _p(per), int(fft_order)))
    return out


def c2c(a, axes=None, forward=True, fct=1.0, nthreads=1):
    """ducc0::c2c (src/ducc0/fft/fft.h:848-850) on a complex array"""
    a = np.ascontiguousarray(a)
    single = a.dtype == np.complex64
    if not single:
        a = a.astype(np.complex128, copy=False)
    ax = np.asarray(range(a.ndim) if axes is None else axes, dtype=np.int64)
    ax = np.where(ax < 0, ax + a.ndim, ax)
    shp = np.asarray(a.shape, dtype=np.int64)
    out = np.empty_like(a)
    _check(lib().ref_c2c(int(single), a.ndim, _p(shp), _p(a), _p(out),
                         _p(ax), len(ax), int(bool(forward)), float(fct),
                         int(nthreads)))
    return out


def _mstart(lmax, mstart):
    if mstart is None:
        ms, idx = [], 0
        for m in range(lmax + 1):
            ms.append(idx)
            idx += lmax - m
        return np.asarray(ms, np.uintp)
    return np.ascontiguousarray(mstart, np.uintp)


_MODES = {"STANDARD": 0, "GRAD_ONLY": 1, "DERIV1": 2}


def sht_synthesis_general(alm, lmax, loc, epsilon, mstart=None, lstride=1,
                          sigma_min=1.1, sigma_max=2.6, nthreads=1, spin=0,
                          mode="STANDARD"):
    """ducc0.sht.synthesis_general (sht.cc:2926-2960): alm (ncomp, nalm) or
    (nalm,) complex64 / complex128 -> map (nmaps, npoints) float32 /
    float64 (nmaps = 1 for spin 0, else 2; spin 0 returns (npoints,))"""
    a = np.ascontiguousarray(alm)
    single = a.dtype == np.complex64
    if not single:
        a = a.astype(np.complex128, copy=False)
    a2 = a.reshape(-1, a.shape[-1])
    ms = _mstart(lmax, mstart)
    lc = np.ascontiguousarray(loc, np.float64)
    nmp = 1 if spin == 0 else 2
    out = np.zeros((nmp, len(lc)), np.float32 if single else np.float64)
    _check(lib().ref_sht_general(int(single), 0, _p(a2), a2.shape[1],
                                 _p(out), lmax, _p(ms), len(ms), lstride,
                                 _p(lc), len(lc), epsilon, sigma_min,
                                 sigma_max, nthreads, spin,
                                 _MODES[mode]))
    return out[0] if spin == 0 else out


def sht_adjoint_synthesis_general(mapv, lmax, loc, epsilon, nalm,
                                  mstart=None, lstride=1, sigma_min=1.1,
                                  sigma_max=2.6, nthreads=1, spin=0,
                                  mode="STANDARD"):
    """ducc0.sht.adjoint_synthesis_general (sht.cc:2962-2993): map
    (nmaps, npoints) or (npoints,) -> alm (ncomp, nalm) (spin 0: (nalm,))"""
    m = np.ascontiguousarray(mapv)
    single = m.dtype == np.float32
    if not single:
        m = m.astype(np.float64, copy=False)
    m2 = m.reshape(-1, m.shape[-1])
    ms = _mstart(lmax, mstart)
    lc = np.ascontiguousarray(loc, np.float64)
    na = 1 if (spin == 0 or mode != "STANDARD") else 2
    out = np.zeros((na, nalm), np.complex64 if single else np.complex128)
    _check(lib().ref_sht_general(int(single), 1, _p(out), nalm, _p(m2), lmax,
                                 _p(ms), len(ms), lstride, _p(lc), len(lc),
                                 epsilon, sigma_min, sigma_max, nthreads,
                                 spin, _MODES[mode]))
    return out[0] if spin == 0 else out


def sht_pseudo_analysis_general(mapv, lmax, loc, spin, maxiter, epsilon,
                                nalm, mstart=None, lstride=1, sigma_min=1.1,
                                sigma_max=2.6, nthreads=1):
    """ducc0.sht.pseudo_analysis_general (sht.cc:3003-3058): map (nmaps,
    npoints) -> (alm (ncomp, nalm), istop, itn, |r|/|b|,
    |A^T r|/(|A||r|))"""
    m = np.ascontiguousarray(mapv)
    single = m.dtype == np.float32
    if not single:
        m = m.astype(np.float64, copy=False)
    m2 = m.reshape(-1, m.shape[-1])
    ms = _mstart(lmax, mstart)
    lc = np.ascontiguousarray(loc, np.float64)
    na = 1 if spin == 0 else 2
    out = np.zeros((na, nalm), np.complex64 if single else np.complex128)
    res = np.zeros(4)
    _check(lib().ref_sht_pseudo_analysis(int(single), _p(out), nalm, _p(m2),
                                         lmax, _p(ms), len(ms), lstride,
                                         _p(lc), len(lc), sigma_min,
                                         sigma_max, nthreads, maxiter,
                                         epsilon, spin, _p(res)))
    return out, int(res[0]), int(res[1]), res[2], res[3]


def wgridder(adjoint, uvw, freq, vis_or_dirty, wgt, mask, npix_x, npix_y,
             pixsize_x, pixsize_y, epsilon, do_wgridding, flip_u=False,
             flip_v=False, flip_w=False, divide_by_n=True, sigma_min=1.1,
             sigma_max=2.6, center_x=0., center_y=0., nthreads=1):
    """ducc0's ms2dirty (adjoint False: vis (nrow, nchan) complex -> dirty)
    or dirty2ms (adjoint True: dirty (npix_x, npix_y) real -> vis),
    wgridder.h:1700-1735; precision from the input's dtype"""
    uvw = np.ascontiguousarray(uvw, np.float64)
    freq = np.ascontiguousarray(freq, np.float64)
    x = np.ascontiguousarray(vis_or_dirty)
    single = x.dtype in (np.complex64, np.float32)
    rt = np.float32 if single else np.float64
    ct_ = np.complex64 if single else np.complex128
    nrow, nchan = uvw.shape[0], freq.shape[0]
    if adjoint:
        dirty = x.astype(rt, copy=False)
        vis = np.zeros((nrow, nchan), ct_)
    else:
        vis = x.astype(ct_, copy=False)
        dirty = np.zeros((npix_x, npix_y), rt)
    w = None if wgt is None else np.ascontiguousarray(wgt, rt)
    m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    _check(lib().ref_wgridder(
        int(single), int(adjoint), _p(uvw), nrow, _p(freq), nchan, _p(vis),
        None if w is None else _p(w), None if m is None else _p(m),
        _p(dirty), npix_x, npix_y, pixsize_x, pixsize_y, epsilon,
        int(do_wgridding), int(flip_u), int(flip_v), int(flip_w),
        int(divide_by_n), sigma_min, sigma_max, center_x, center_y,
        nthreads))
    return vis if adjoint else dirty


class Plan:
    """ducc0.nufft.plan (nufft_pymod.cc:139-261) over the reference's Nufft
    class: coordinates and values in one precision (float32 / complex64
    when `single`)."""

    def __init__(self, nu2u, coord, grid_shape, epsilon, single,
                 nthreads=1, sigma_min=1.1, sigma_max=2.6,
                 periodicity=2 * np.pi, fft_order=False):
        self.single = bool(single)
        self.c = np.ascontiguousarray(coord, np.float32 if single
                                      else np.float64)
        self.shape = tuple(int(x) for x in grid_shape)
        self.M = self.c.shape[0]
        per = np.ascontiguousarray(np.broadcast_to(
            np.asarray(periodicity, float), (len(self.shape),)))
        h = ct.c_void_p()
        _check(lib().ref_plan_create(
            ct.byref(h), int(bool(nu2u)), len(self.shape), self.M,
            _p(self.c), int(self.single),
            _p(np.asarray(self.shape, np.uintp)), float(epsilon),
            int(nthreads), float(sigma_min), float(sigma_max), _p(per),
            int(bool(fft_order))))
        self.h = h

    def _vt(self):
        return np.complex64 if self.single else np.complex128

    def nu2u(self, points, forward=True):
        v = np.ascontiguousarray(points, self._vt())
        out = np.zeros(self.shape, self._vt())
        _check(lib().ref_plan_nu2u(self.h, int(bool(forward)), _p(v),
                                   _p(out)))
        return out

    def u2nu(self, grid, forward=True):
        g = np.ascontiguousarray(grid, self._vt())
        out = np.zeros(self.M, self._vt())
        _check(lib().ref_plan_u2nu(self.h, int(bool(forward)), _p(g),
                                   _p(out)))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().ref_plan_destroy(self.h)
            self.h = None
