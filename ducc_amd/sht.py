"""ducc0.sht general synthesis on the GPU: the Python surface of
ducc0.sht.synthesis_general / adjoint_synthesis_general /
pseudo_analysis_general (python/sht_pymod.cc:805-927, 2299-2301),
keyword-only, with ducc0's defaults.

Arrays may be numpy (copied to and from the device) or torch tensors on the
GPU (used in place, ordered on torch's current stream).  a_lm are complex
(ncomp, nalm), the map real (nmaps, npoints) of the same precision, loc
(npoints, 2) float64 (theta, phi).  nmaps is 1 for spin 0, else 2 (Q, U);
ncomp is 1 for spin 0 and in the modes GRAD_ONLY (E alone) and DERIV1 (the
gradient of a spin-0 field, spin 1), else 2 (E, B)."""
import ctypes as ct

import numpy as np

from . import _lib
from ._lib import _is_torch


def _mstart(lmax, mmax, mstart):
    """ducc0's a_lm layout (sht_pymod.cc:124-140): by default m-major, a_lm
    at mstart[m] + l with mstart[m] = m * (2 * lmax + 1 - m) / 2"""
    if mstart is None:
        mmax = lmax if mmax is None else int(mmax)
        if mmax > lmax:
            raise RuntimeError("mmax>lmax")
        ms = np.zeros(mmax + 1, np.int64)
        idx = 0
        for m in range(mmax + 1):
            ms[m] = idx
            idx += lmax - m
        return ms
    ms = np.ascontiguousarray(mstart, np.int64)
    if mmax is not None and int(mmax) + 1 != len(ms):
        raise RuntimeError("mmax and mstart size mismatch")
    if len(ms) > lmax + 1:
        raise RuntimeError("mmax>lmax")
    return ms


def _min_almdim(lmax, ms, lstride):
    res = 0
    for m in range(len(ms)):
        lo = int(ms[m]) + m * lstride
        hi = int(ms[m]) + lmax * lstride
        if lo < 0 or hi < 0:
            raise RuntimeError("impossible a_lm memory layout")
        res = max(res, lo, hi)
    return res + 1


_MODES = {"STANDARD": 0, "GRAD_ONLY": 1, "DERIV1": 2}


def _mode(mode):
    """sht_pymod.cc:50-56"""
    if mode not in _MODES:
        raise RuntimeError("unknown SHT mode")
    return _MODES[mode]


def _ncomp(spin, mode):
    """(a_lm components, map components) (sht.cc:2931-2935)"""
    nalm = 1 if (spin == 0 or mode != "STANDARD") else 2
    return nalm, (1 if spin == 0 else 2)


def _stream_opts(*arrays):
    if any(_is_torch(a) and a.is_cuda for a in arrays):
        import torch
        return _lib.options(stream=torch.cuda.current_stream().cuda_stream)
    return _lib.options()


def synthesis_general(*, alm, spin, lmax, loc, epsilon, mstart=None,
                      lstride=1, mmax=None, nthreads=1, map=None,
                      sigma_min=1.1, sigma_max=2.6, mode="STANDARD",
                      verbose=False):
    """Values of the real field sum_{l,m} a_lm Y_lm (spin 0), or of the spin
    field Q + i U = -sum_{l,m} (E_lm + i B_lm) sY_lm, at the points loc[i] =
    (theta_i, phi_i) (sht.cc:2926-2960).  Returns `map` (the same object
    when given), shape (nmaps, npoints)."""
    _lib.require_gpu()
    mo = _mode(mode)
    _, nmaps = _ncomp(spin, mode)
    ms = _mstart(lmax, mmax, mstart)
    _min_almdim(lmax, ms, lstride)
    npt = loc.shape[0]
    if map is None:
        if _is_torch(alm):
            import torch
            rdt = torch.float64 if alm.dtype == torch.complex128 \
                else torch.float32
            map = torch.empty((nmaps, npt), dtype=rdt, device=alm.device)
        else:
            a = np.asarray(alm)
            rdt = np.float64 if a.dtype == np.complex128 else np.float32
            map = np.empty((nmaps, npt), dtype=rdt)
    a, ka = _lib.describe(alm)
    m_, km = _lib.describe(map)
    lc, kl = _lib.describe(loc)
    opt = _stream_opts(alm, map, loc)
    _lib.check(_lib.lib().dmx_sht_synthesis_general(
        ct.byref(a), ct.byref(m_), int(spin), int(lmax), ms.ctypes.data,
        len(ms), int(lstride), ct.byref(lc), float(epsilon),
        float(sigma_min), float(sigma_max), int(nthreads), mo,
        int(bool(verbose)), ct.byref(opt)))
    return map


def adjoint_synthesis_general(*, map, spin, lmax, loc, epsilon, mstart=None,
                              lstride=1, mmax=None, nthreads=1, alm=None,
                              sigma_min=1.1, sigma_max=2.6,
                              mode="STANDARD", verbose=False):
    """The adjoint of synthesis_general (sht.cc:2962-2993): a_lm from the
    values `map` (nmaps, npoints) at loc.  Returns `alm` (the same object
    when given), shape (ncomp, nalm)."""
    _lib.require_gpu()
    mo = _mode(mode)
    ncomp, _ = _ncomp(spin, mode)
    ms = _mstart(lmax, mmax, mstart)
    nalm = _min_almdim(lmax, ms, lstride)
    if alm is None:
        if _is_torch(map):
            import torch
            cdt = torch.complex128 if map.dtype == torch.float64 \
                else torch.complex64
            alm = torch.zeros((ncomp, nalm), dtype=cdt, device=map.device)
        else:
            cdt = np.complex128 if np.asarray(map).dtype == np.float64 \
                else np.complex64
            alm = np.zeros((ncomp, nalm), dtype=cdt)
    a, ka = _lib.describe(alm)
    m_, km = _lib.describe(map)
    lc, kl = _lib.describe(loc)
    opt = _stream_opts(alm, map, loc)
    _lib.check(_lib.lib().dmx_sht_adjoint_synthesis_general(
        ct.byref(a), ct.byref(m_), int(spin), int(lmax), ms.ctypes.data,
        len(ms), int(lstride), ct.byref(lc), float(epsilon),
        float(sigma_min), float(sigma_max), int(nthreads), mo,
        int(bool(verbose)), ct.byref(opt)))
    return alm


def pseudo_analysis_general(*, lmax, map, loc, spin, nthreads, maxiter,
                            epsilon=1e-5, sigma_min=1.1, sigma_max=2.6,
                            mstart=None, lstride=1, mmax=None, alm=None,
                            verbose=False):
    """a_lm whose synthesis_general best fits `map` at loc (least squares),
    by LSMR (sht.cc:3003-3058, sht_pymod.cc:885-927: atol 1e-14
    sqrt(npoints), btol epsilon, at most maxiter iterations, the NUFFTs at
    0.1 epsilon).  Returns (alm, istop, itn, |r| / |b|,
    |A^T r| / (|A| |r|)) with ducc0's meaning of istop (1: solution found,
    2: least-squares solution found, 3: condition too large, 7: maxiter
    reached); `alm` is the same object when given."""
    _lib.require_gpu()
    ncomp, _ = _ncomp(spin, "STANDARD")
    ms = _mstart(lmax, mmax, mstart)
    nalm = _min_almdim(lmax, ms, lstride)
    if alm is None:
        if _is_torch(map):
            import torch
            cdt = torch.complex128 if map.dtype == torch.float64 \
                else torch.complex64
            alm = torch.zeros((ncomp, nalm), dtype=cdt, device=map.device)
        else:
            cdt = np.complex128 if np.asarray(map).dtype == np.float64 \
                else np.complex64
            alm = np.zeros((ncomp, nalm), dtype=cdt)
    a, ka = _lib.describe(alm)
    m_, km = _lib.describe(map)
    lc, kl = _lib.describe(loc)
    opt = _stream_opts(alm, map, loc)
    istop, itn = ct.c_size_t(0), ct.c_size_t(0)
    rn, arn = ct.c_double(0), ct.c_double(0)
    _lib.check(_lib.lib().dmx_sht_pseudo_analysis_general(
        ct.byref(a), ct.byref(m_), int(spin), int(lmax), ms.ctypes.data,
        len(ms), int(lstride), ct.byref(lc), float(sigma_min),
        float(sigma_max), int(nthreads), int(maxiter), float(epsilon),
        int(bool(verbose)), ct.byref(istop), ct.byref(itn), ct.byref(rn),
        ct.byref(arn), ct.byref(opt)))
    return alm, istop.value, itn.value, rn.value, arn.value
