"""The C-ABI boundary without a GPU: the library loads, exports every symbol
include/ducc_amd.h declares, and validates arguments on the host with the
reference's error messages before touching a device."""
import ctypes as ct
import json
import os
import re

import numpy as np
import pytest

from ducc_amd import _lib
from oracle import core as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "ducc_amd.h")).read()
    return sorted(set(re.findall(r"\b(dmx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    L = ct.CDLL(_lib.LIBPATH)
    syms = header_symbols()
    assert len(syms) >= 14
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(_lib.exported_names())


@pytest.mark.parametrize("ndim", [1, 2, 3])
@pytest.mark.parametrize("single", [True, False])
def test_best_epsilon_matches_oracle(ndim, single):
    from ducc_amd import nufft
    for smin, smax in ((1.1, 2.6), (1.2, 2.51), (1.5, 2.0)):
        a = nufft.bestEpsilon(ndim=ndim, singleprec=single, sigma_min=smin,
                              sigma_max=smax)
        assert a == O.best_epsilon(ndim, single, smin, smax)


def test_best_epsilon_errors():
    from ducc_amd import nufft
    with pytest.raises(RuntimeError, match="bad dimensionality"):
        nufft.bestEpsilon(ndim=4, singleprec=False)
    with pytest.raises(RuntimeError, match="no appropriate kernel found"):
        nufft.bestEpsilon(ndim=2, singleprec=False, sigma_min=3.0,
                          sigma_max=3.5)


def _call_u2nu(grid, coord, eps=1e-5, per=None):
    g, kg = _lib.describe(grid)
    c, kc = _lib.describe(coord)
    nd = coord.shape[1] if coord.ndim == 2 else 1
    out = np.zeros(coord.shape[0], dtype=np.complex128)
    o, ko = _lib.describe(out)
    per = np.full(max(nd, 1), 2 * np.pi) if per is None else per
    opt = _lib.options()
    rc = _lib.lib().dmx_u2nu(ct.byref(g), ct.byref(c), 1, eps, 1, ct.byref(o),
                             0, 1.2, 2.51, per.ctypes.data, 0, ct.byref(opt),
                             None)
    return rc, _lib.lib().dmx_last_error().decode()


def test_host_validation_messages():
    """Argument errors are raised on the host (no device needed) with the
    reference's messages (nufft.h:342-343, 1904-1905;
    gridding_kernel.h:442; nufft_pymod.cc:93)."""
    c = np.zeros((10, 2))
    rc, msg = _call_u2nu(np.zeros((4, 4), np.complex64).astype(np.float32), c)
    assert rc and "not yet supported" in msg
    rc, msg = _call_u2nu(np.zeros((4, 4, 4), np.complex128), c)
    assert rc and "dimensionality mismatch" in msg
    rc, msg = _call_u2nu(np.zeros((4, 4), np.complex128), c, eps=1e-30)
    assert rc and "no appropriate kernel found" in msg
    rc, msg = _call_u2nu(np.zeros((4, 4), np.complex128), c, eps=-1.0)
    assert rc and "epsilon must be positive" in msg
    rc, msg = _call_u2nu(np.zeros((2, 2, 2, 2), np.complex128),
                         np.zeros((10, 4)))
    assert rc and "transform must be 1D/2D/3D" in msg


def _choose(shape, M, eps, single, gridding, smin, smax):
    """The library's analytic choice through dmx_choose (host only)."""
    shp = np.asarray(shape, np.int64)
    opt = _lib.options(precision="fp32" if single else "fp64")
    inf = _lib.DmxPlanInfo()
    rc = _lib.lib().dmx_choose(len(shape), shp.ctypes.data, int(M), float(eps),
                               int(not single), int(bool(gridding)),
                               float(smin), float(smax), ct.byref(opt),
                               ct.byref(inf))
    if rc:
        return -1, None, _lib.lib().dmx_last_error().decode()
    return inf.kernel_idx, tuple(inf.nover[:len(shape)]), ""


def test_params_match_ducc0_golden():
    """dmx_choose (no GPU needed) returns ducc0's kernel row and oversampled
    grid for every golden request that ducc0 ran with 8 threads at vlen 8
    (the host model the library restates; tests/golden/params.json was
    written by the reference built from its sources)."""
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "params.json")))
    assert g["vlen_single"] == 8 and g["vlen_double"] == 8
    n = 0
    for r in g["requests"]:
        if r["nthreads"] != 8:
            continue
        k, nov, msg = _choose(r["shape"], r["M"], r["eps"], r["single"],
                              r["gridding"], r["sigma_min"], r["sigma_max"])
        assert k == r["kidx"], (r, msg)
        if k >= 0:
            assert list(nov) == r["nover"], r
        else:
            assert "no appropriate kernel found" in msg
        n += 1
    assert n >= 20


def test_params_match_oracle_restatement():
    """dmx_choose equals the oracle's find_params (itself pinned to the
    golden requests above and to live ducc0) on 400 random requests over
    1-3 D, both precisions, both transform types and random sigma ranges."""
    rng = np.random.default_rng(7)
    nk = 0
    for _ in range(400):
        nd = int(rng.integers(1, 4))
        hi = {1: 200000, 2: 5000, 3: 400}[nd]
        shape = tuple(int(x) for x in np.exp(rng.uniform(np.log(2),
                                                         np.log(hi), nd)))
        single = bool(rng.integers(0, 2))
        gridding = bool(rng.integers(0, 2))
        eps = 10 ** rng.uniform(-7.5 if single else -14.5, -1)
        M = int(10 ** rng.uniform(0, 8))
        smin = float(rng.choice([1.1, 1.2, 1.25, 1.5]))
        smax = float(rng.choice([2.0, 2.3, 2.51, 2.6]))
        k_o, nov_o = O.find_params(eps, shape, M, gridding, single, vlen=8,
                                   nthreads=8, sigma_min=smin,
                                   sigma_max=smax)
        k, nov, msg = _choose(shape, M, eps, single, gridding, smin, smax)
        assert k == k_o, (nd, shape, single, gridding, eps, M, smin, smax,
                          msg)
        if k >= 0:
            assert tuple(nov) == tuple(nov_o)
            nk += 1
    assert nk > 300


def test_params_match_live_ducc0():
    """dmx_choose equals ducc0's own findNufftParameters (the reference
    built from its sources into oracle/_ref) on 300 random requests at 8
    threads; skipped only where that build is absent."""
    from oracle import ref as R
    if not R.available():
        pytest.skip("oracle/_ref (ducc0 built from the reference) is absent")
    rng = np.random.default_rng(11)
    for _ in range(300):
        nd = int(rng.integers(1, 4))
        hi = {1: 100000, 2: 4096, 3: 300}[nd]
        shape = tuple(int(x) for x in np.exp(rng.uniform(np.log(4),
                                                         np.log(hi), nd)))
        single = bool(rng.integers(0, 2))
        gridding = bool(rng.integers(0, 2))
        eps = 10 ** rng.uniform(-6 if single else -13, -1.5)
        M = int(10 ** rng.uniform(1, 7.5))
        k_r, nov_r = R.find_params(eps, shape, M, gridding, single, 8,
                                   sigma_min=1.2, sigma_max=2.51)
        k, nov, msg = _choose(shape, M, eps, single, gridding, 1.2, 2.51)
        assert k == k_r, (nd, shape, single, gridding, eps, M, msg)
        if k >= 0:
            assert tuple(nov) == tuple(nov_r)
