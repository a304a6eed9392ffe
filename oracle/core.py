"""ctypes wrapper over oracle.c, the CPU restatement of ducc0's NUFFT path.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Each wrapper names the
reference lines its C function follows.
"""
import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
DB_CSV = os.path.join(os.path.dirname(HERE), "ducc_amd", "data",
                      "kernel_db.csv")

_lib = None
_db = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def load_db():
    """KernelDB rows (src/ducc0/math/gridding_kernel.cc:78-1537) as arrays."""
    global _db
    if _db is None:
        a = np.loadtxt(DB_CSV, delimiter=",", comments="#")
        _db = dict(W=a[:, 0].astype(np.int32), ofactor=a[:, 1].copy(),
                   epsilon=a[:, 2].copy(), beta=a[:, 3].copy(),
                   e0=a[:, 4].copy(), ndim=a[:, 5].astype(np.int32),
                   single=a[:, 6].astype(np.int32))
    return _db


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ct.CDLL(LIB)
        P, I, D, S = ct.c_void_p, ct.c_int, ct.c_double, ct.c_size_t
        L.orc_good_size_cmplx.restype = S
        L.orc_good_size_cmplx.argtypes = [S]
        L.orc_best_epsilon.restype = D
        L.orc_best_epsilon.argtypes = [I, I, D, D]
        L.orc_find_params.restype = I
        L.orc_find_params.argtypes = [D, D, D, I, P, S, I, I, I, S, D, P]
        L.orc_available_kernels.restype = I
        L.orc_available_kernels.argtypes = [D, I, I, D, D, P]
        L.orc_kernel_coeffs.argtypes = [I, I, D, D, P]
        L.orc_gl_nodes.argtypes = [I, P, P]
        L.orc_corfac.argtypes = [I, D, D, S, D, P]
        L.orc_esk.restype = D
        L.orc_esk.argtypes = [D, D, D]
        L.orc_fold.argtypes = [I, I, P, P, I, S, P, P, P]
        L.orc_keys.argtypes = [I, I, P, P, I, I, I, S, P, P]
        L.orc_sort_stable.argtypes = [S, P, ct.c_uint32, P]
        L.orc_fft_c2c.argtypes = [I, P, P, I]
        L.orc_nu2u.restype = I
        L.orc_nu2u.argtypes = [I, S, P, P, I, D, D, D, P, I, P, I, I, S, I,
                               P, P, P]
        L.orc_u2nu.restype = I
        L.orc_u2nu.argtypes = L.orc_nu2u.argtypes
        L.orc_nudft_type1.argtypes = [I, S, P, P, I, P, I, P, S, P, P]
        L.orc_nudft_type2.argtypes = [I, S, P, P, I, P, I, P, S, P, P]
        L.orc_l2error.restype = D
        L.orc_l2error.argtypes = [S, P, P]
        L.orc_db_row.argtypes = [I, P, P, P, P, P]
        db = load_db()
        L.orc_set_db(ct.c_int(len(db["W"])),
                     *[db[k].ctypes.data_as(P) for k in
                       ("W", "ofactor", "epsilon", "beta", "e0", "ndim",
                        "single")])
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ct.c_void_p) if a is not None else None


def _sz(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.uintp))


def good_size(n):
    """fft.h:93-119"""
    return int(lib().orc_good_size_cmplx(n))


def best_epsilon(ndim, single, sigma_min=1.1, sigma_max=2.6):
    """gridding_kernel.cc:1562-1574"""
    return lib().orc_best_epsilon(ndim, int(single), sigma_min, sigma_max)


def available_kernels(eps, ndim, single, sigma_min, sigma_max):
    """gridding_kernel.h:420-444"""
    out = np.zeros(20, np.int32)
    n = lib().orc_available_kernels(eps, ndim, int(single), sigma_min,
                                    sigma_max, _p(out))
    return out[:n].copy()


def find_params(eps, shape, npoints, gridding, single, vlen=8, nthreads=8,
                sigma_min=1.2, sigma_max=2.51):
    """nufft.h:155-208 -> (KernelDB row, oversampled shape)"""
    nover = np.zeros(3, np.uintp)
    dims = _sz(shape)
    k = lib().orc_find_params(eps, sigma_min, sigma_max, len(shape), _p(dims),
                              npoints, int(gridding), int(single), vlen,
                              nthreads, 1.0, _p(nover))
    return k, tuple(int(x) for x in nover[:len(shape)])


def db_row(k):
    db = load_db()
    return dict(W=int(db["W"][k]), ofactor=float(db["ofactor"][k]),
                epsilon=float(db["epsilon"][k]), beta=float(db["beta"][k]),
                e0=float(db["e0"][k]))


def esk(v, beta, e0):
    """gridding_kernel.cc:1539-1544 (beta already times W)"""
    return lib().orc_esk(v, beta, e0)


def kernel_coeffs(W, beta_row, e0, D=None):
    """gridding_kernel.cc:28-76, :1552-1560 -> (D+1, W) array, row j = x^(D-j)"""
    D = W + 3 if D is None else D
    c = np.zeros((D + 1) * W)
    lib().orc_kernel_coeffs(W, D, beta_row * W, e0, _p(c))
    return c.reshape(D + 1, W)


def gl_nodes(n):
    """gl_integrator.cc:38-77,220-235: non-negative nodes and weights"""
    m = (n + 1) // 2
    x = np.zeros(m)
    w = np.zeros(m)
    lib().orc_gl_nodes(n, _p(x), _p(w))
    return x, w


def corfac(W, beta_row, e0, n, dx):
    """gridding_kernel.h:68-129"""
    out = np.zeros(n)
    lib().orc_corfac(W, beta_row * W, e0, n, dx, _p(out))
    return out


def fold(coords, W, nover, period, coord_double=True):
    """nufft.h:285-298 -> (i0 int32 (M,d), frac float64 (M,d))"""
    c = np.ascontiguousarray(coords, dtype=np.float64)
    M, d = c.shape
    i0 = np.zeros((M, d), np.int32)
    fr = np.zeros((M, d))
    per = np.ascontiguousarray(np.broadcast_to(period, (d,)), np.float64)
    lib().orc_fold(d, W, _p(_sz(nover)), _p(per), int(coord_double), M,
                   _p(c), _p(i0), _p(fr))
    return i0, fr


def tile_keys(coords, W, nover, period, log2tile, lsq2=None,
              coord_double=True):
    """nufft.h:301-320, 892-907, 1364-1381, 1862-1892"""
    c = np.ascontiguousarray(coords, dtype=np.float64)
    M, d = c.shape
    keys = np.zeros(M, np.uint32)
    per = np.ascontiguousarray(np.broadcast_to(period, (d,)), np.float64)
    lsq2 = log2tile if lsq2 is None else lsq2
    lib().orc_keys(d, W, _p(_sz(nover)), _p(per), int(coord_double),
                   log2tile, lsq2, M, _p(c), _p(keys))
    return keys


def sort_stable(keys):
    """bucket_sort.h:170-354 ordering"""
    k = np.ascontiguousarray(keys, np.uint32)
    idx = np.zeros(len(k), np.uint32)
    mx = int(k.max()) if len(k) else 0
    lib().orc_sort_stable(len(k), _p(k), mx, _p(idx))
    return idx


def fft_c2c(a, forward):
    """unnormalised c2c over all axes (the role of fft.h:848-850)"""
    x = np.ascontiguousarray(a, np.complex128).copy()
    lib().orc_fft_c2c(x.ndim, _p(_sz(x.shape)), _p(x), int(forward))
    return x


def _per(period, d):
    return np.ascontiguousarray(np.broadcast_to(np.asarray(period, float),
                                                (d,)), np.float64)


def nu2u(coords, points, shape, forward, eps, single=False, vlen=8,
         nthreads=8, sigma_min=1.2, sigma_max=2.51, period=2 * np.pi,
         fft_order=False, coord_double=True, info=None):
    """Type 1, nufft.h:836-862/1274-1315/1731-1794 (computed in float64)"""
    c = np.ascontiguousarray(coords, np.float64)
    p = np.ascontiguousarray(points, np.complex128)
    M, d = c.shape
    out = np.zeros(shape, np.complex128)
    kidx = np.zeros(1, np.int32)
    nover = np.zeros(3, np.uintp)
    rc = lib().orc_nu2u(d, M, _p(c), _p(p), int(forward), eps, sigma_min,
                        sigma_max, _p(_per(period, d)), int(fft_order),
                        _p(_sz(shape)), int(single), vlen, nthreads,
                        int(coord_double), _p(out), _p(kidx), _p(nover))
    if rc:
        raise RuntimeError("no appropriate kernel found")
    if info is not None:
        info.update(kidx=int(kidx[0]), nover=tuple(int(x) for x in nover[:d]))
    return out


def u2nu(coords, grid, forward, eps, single=False, vlen=8, nthreads=8,
         sigma_min=1.2, sigma_max=2.51, period=2 * np.pi, fft_order=False,
         coord_double=True, info=None):
    """Type 2, nufft.h:864-890/1317-1362/1796-1860 (computed in float64)"""
    c = np.ascontiguousarray(coords, np.float64)
    g = np.ascontiguousarray(grid, np.complex128)
    M, d = c.shape
    out = np.zeros(M, np.complex128)
    kidx = np.zeros(1, np.int32)
    nover = np.zeros(3, np.uintp)
    rc = lib().orc_u2nu(d, M, _p(c), _p(g), int(forward), eps, sigma_min,
                        sigma_max, _p(_per(period, d)), int(fft_order),
                        _p(_sz(g.shape)), int(single), vlen, nthreads,
                        int(coord_double), _p(out), _p(kidx), _p(nover))
    if rc:
        raise RuntimeError("no appropriate kernel found")
    if info is not None:
        info.update(kidx=int(kidx[0]), nover=tuple(int(x) for x in nover[:d]))
    return out


def nudft_type1(coords, points, shape, forward, period=2 * np.pi,
                fft_order=False, modes=None):
    """Brute-force type 1 (python/test/test_nufft.py:32-43).  `modes`: flat
    user indices to evaluate (all when None)."""
    c = np.ascontiguousarray(coords, np.float64)
    p = np.ascontiguousarray(points, np.complex128)
    M, d = c.shape
    if modes is None:
        out = np.zeros(int(np.prod(shape)), np.complex128)
        m = None
        nm = out.size
    else:
        m = np.ascontiguousarray(modes, np.int64)
        out = np.zeros(len(m), np.complex128)
        nm = len(m)
    lib().orc_nudft_type1(d, M, _p(c), _p(p), int(forward),
                          _p(_per(period, d)), int(fft_order),
                          _p(_sz(shape)), nm, _p(m), _p(out))
    return out.reshape(shape) if modes is None else out


def nudft_type2(coords, grid, forward, period=2 * np.pi, fft_order=False,
                pidx=None):
    """Brute-force type 2 (adjoint of test_nufft.py:32-43's sum)."""
    c = np.ascontiguousarray(coords, np.float64)
    g = np.ascontiguousarray(grid, np.complex128)
    M, d = c.shape
    if pidx is None:
        out = np.zeros(M, np.complex128)
        pi_ = None
        n = M
    else:
        pi_ = np.ascontiguousarray(pidx, np.int64)
        out = np.zeros(len(pi_), np.complex128)
        n = len(pi_)
    lib().orc_nudft_type2(d, M, _p(c), _p(g), int(forward),
                          _p(_per(period, d)), int(fft_order),
                          _p(_sz(g.shape)), n, _p(pi_), _p(out))
    return out


def l2error(a, b):
    """python/misc_pymod.cc:129-170"""
    x = np.ascontiguousarray(a, np.complex128).ravel()
    y = np.ascontiguousarray(b, np.complex128).ravel()
    return lib().orc_l2error(x.size, _p(x), _p(y))
