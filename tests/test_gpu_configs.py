"""Every workload configuration at its full size, against the FP64 direct
sum (the GPU NUDFT, itself pinned to the oracle's sums in
test_gpu_nudft.py), the oracle's CPU sums where they finish in seconds, and
ducc0 itself (built from the reference sources into oracle/_ref):

  C1  2-D type 1, complex128, M = 1e5 uniform, 256^2, eps 1e-9
  C2  2-D type 1 and type 2 on the clustered headline geometry: M = 1e7
      points in 8 Gaussian clusters, 2048^2, eps 1e-5, complex64 (FP32
      path) and the plain FP64 path
  C3  3-D type 1, M = 5e7 uniform, 256^3, eps 1e-4 / 1e-6 (complex64) and
      1e-9 (complex128, FP64 path), 10,000 sampled modes
  C4  differentiable 2-D type 2 in complex64: 512^2, M = 2e6, eps 1e-6,
      gradients against FP64 central differences (h = 1e-6)
  C5  the learned chooser over 1,000 requests of the mixed distribution

Error bounds follow the reference's tests (python/test/test_nufft.py:79,
140, 203: 4/2/1 eps in 1/2/3-D) unless stated.  Inputs are regenerated
from numpy's PCG64 with the seeds 101-105."""
import math
import os

import numpy as np
import pytest
import torch

from oracle import core as O

pytestmark = pytest.mark.gpu


def _gauss(rng, shape, dtype=np.complex128):
    z = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return (z * np.sqrt(0.5)).astype(dtype)


def _clustered(rng, M, d=2, ncl=8, sig=0.3):
    cen = rng.uniform(-np.pi, np.pi, (ncl, d))
    lab = rng.integers(0, ncl, M)
    x = cen[lab] + rng.normal(0.0, sig, (M, d))
    return (x + np.pi) % (2 * np.pi) - np.pi


def _ref():
    from oracle import ref as R
    return R if R.available() else None


def _l2(a, b):
    return O.l2error(np.asarray(a, np.complex128), np.asarray(b, np.complex128))


def test_c1_full(gpu):
    from ducc_amd import nufft
    rng = np.random.default_rng(101)
    M, n, eps = 100_000, 256, 1e-9
    c = rng.uniform(-np.pi, np.pi, (M, 2))
    p = _gauss(rng, M)
    out = torch.empty((n, n), dtype=torch.complex128, device="cuda")
    info = {}
    nufft.nu2u(points=torch.from_numpy(p).cuda(),
               coord=torch.from_numpy(c).cuda(), forward=True, epsilon=eps,
               out=out, info=info)
    got = out.cpu().numpy()
    assert info["fp64"] and info["W"] == 11 and info["nover"] == (462, 462)
    # the oracle's CPU sum on 300 modes, the GPU sum on 3,000
    sel = rng.choice(n * n, 300, replace=False)
    ref = O.nudft_type1(c, p, (n, n), True, modes=sel)
    assert _l2(got.ravel()[sel], ref) <= 2 * eps
    sel2 = rng.choice(n * n, 3000, replace=False)
    ref2 = nufft.nudft_direct(1, c, p, (n, n), True, sel2)
    assert _l2(got.ravel()[sel2], ref2) <= 2 * eps
    R = _ref()
    if R is not None:
        r = R.nu2u(p, c, True, eps, (n, n), nthreads=os.cpu_count() or 1)
        assert _l2(got, r) <= 2 * eps


@pytest.fixture(scope="module")
def c2_inputs():
    rng = np.random.default_rng(102)
    M, n = 10_000_000, 2048
    c = _clustered(rng, M)
    p = _gauss(rng, M, np.complex64)
    g = _gauss(rng, (n, n), np.complex64)
    return rng, c, p, g


@pytest.mark.parametrize("type_", [1, 2])
def test_c2_headline_full(gpu, c2_inputs, type_):
    """FP32 (complex64) and plain FP64 (complex128, ducc0's analytic
    parameters) on the headline geometry: each within eps of the direct sum
    on 256 sampled modes (type 1) or points (type 2), within eps of each
    other on the whole output, and the FP64 path within eps of ducc0's own
    FP64 transform of the same inputs."""
    from ducc_amd import nufft
    rng0, c, p, g = c2_inputs
    rng = np.random.default_rng(1020 + type_)
    M, n, eps = len(p), g.shape[0], 1e-5
    ct = torch.from_numpy(c).cuda()
    res = {}
    for prec, dt in (("auto", torch.complex64), ("fp64", torch.complex128)):
        info = {}
        if type_ == 1:
            out = torch.empty((n, n), dtype=dt, device="cuda")
            nufft.nu2u(points=torch.from_numpy(p).cuda().to(dt), coord=ct,
                       forward=True, epsilon=eps, out=out, precision=prec,
                       info=info)
        else:
            out = nufft.u2nu(grid=torch.from_numpy(g).cuda().to(dt), coord=ct,
                             forward=False, epsilon=eps, precision=prec,
                             info=info)
        res[prec] = (out.cpu().numpy().astype(np.complex128), info)
        del out
    assert not res["auto"][1]["fp64"] and res["fp64"][1]["fp64"]
    if type_ == 1:
        sel = rng.choice(n * n, 256, replace=False)
        d = nufft.nudft_direct(1, c, p.astype(np.complex128), (n, n), True,
                               sel)
        take = lambda a: a.ravel()[sel]
    else:
        sel = rng.choice(M, 256, replace=False)
        d = nufft.nudft_direct(2, c, g.astype(np.complex128), (n, n), False,
                               sel)
        take = lambda a: a[sel]
    for prec in ("auto", "fp64"):
        e = _l2(take(res[prec][0]), d)
        assert e <= eps, (prec, e)
    assert _l2(res["auto"][0], res["fp64"][0]) <= eps
    R = _ref()
    if R is not None:
        nth = os.cpu_count() or 1
        if type_ == 1:
            r = R.nu2u(p.astype(np.complex128), c, True, eps, (n, n),
                       nthreads=nth)
        else:
            r = R.u2nu(g.astype(np.complex128), c, False, eps, nthreads=nth)
        assert _l2(res["fp64"][0], r) <= eps
        assert _l2(res["auto"][0], r) <= eps


@pytest.fixture(scope="module")
def c3_inputs():
    rng = np.random.default_rng(103)
    M, n = 50_000_000, 256
    c = rng.uniform(-np.pi, np.pi, (M, 3))
    # complex64-representable values: the FP64 run (complex128) and the
    # complex64 runs transform the same numbers, so one reference serves
    p = _gauss(rng, M, np.complex64)
    sel = rng.choice(n ** 3, 10_000, replace=False)
    from ducc_amd import nufft
    ref = nufft.nudft_direct(1, c, p.astype(np.complex128), (n, n, n), True,
                             sel)
    return c, p, sel, ref


@pytest.mark.parametrize("eps,dt,fp64", [(1e-4, torch.complex64, False),
                                         (1e-6, torch.complex64, False),
                                         (1e-9, torch.complex128, True)])
def test_c3_full(gpu, c3_inputs, eps, dt, fp64):
    from ducc_amd import nufft
    c, p, sel, ref = c3_inputs
    n = 256
    info = {}
    out = torch.empty((n, n, n), dtype=dt, device="cuda")
    nufft.nu2u(points=torch.from_numpy(p).cuda().to(dt),
               coord=torch.from_numpy(c).cuda(), forward=True, epsilon=eps,
               out=out, info=info)
    assert info["fp64"] == fp64
    got = out.cpu().numpy().ravel()[sel]
    del out
    e = _l2(got, ref)
    assert e <= eps, e


def test_c4_complex64_gradients(gpu):
    """The complex64 autograd path at full size: gradients of
    L = sum |u2nu(g, x) - t|^2 with respect to the grid and the point
    coordinates against FP64 central differences (h = 1e-6, computed with
    the complex128 path) on 32 grid entries (real and imaginary parts) and
    32 points (both axes), and against the complex128 gradients as a
    whole."""
    from ducc_amd import nufft
    from ducc_amd import torch_nufft as tn
    rng = np.random.default_rng(104)
    M, n, eps, h = 2_000_000, 512, 1e-6, 1e-6
    c0 = rng.uniform(-np.pi, np.pi, (M, 2))
    g0 = _gauss(rng, (n, n))
    t0 = _gauss(rng, M)
    grads = {}
    for dt in (torch.complex64, torch.complex128):
        gg = torch.tensor(g0, device="cuda", dtype=dt, requires_grad=True)
        cg = torch.tensor(c0, device="cuda", requires_grad=True)
        tg = torch.tensor(t0, device="cuda", dtype=dt)
        L = (tn.u2nu(gg, cg, True, eps=eps) - tg).abs().pow(2).sum()
        L.backward()
        grads[dt] = (gg.grad.cpu().numpy().astype(np.complex128),
                     cg.grad.cpu().numpy())
    g32, c32 = grads[torch.complex64]
    g64, c64 = grads[torch.complex128]
    assert np.linalg.norm(g32 - g64) <= 1e-4 * np.linalg.norm(g64)
    assert np.linalg.norm(c32 - c64) <= 1e-3 * np.linalg.norm(c64)

    tg = torch.tensor(t0, device="cuda")
    g_t = torch.tensor(g0, device="cuda")
    c_t = torch.tensor(c0, device="cuda")

    def dloss(gp, cp, gm, cm):
        op = nufft.u2nu(grid=gp, coord=cp, forward=True, epsilon=eps)
        om = nufft.u2nu(grid=gm, coord=cm, forward=True, epsilon=eps)
        return float(torch.sum((op - tg).abs() ** 2 -
                                (om - tg).abs() ** 2).item()) / (2 * h)
    errs = []
    for q in rng.choice(n * n, 32, replace=False):
        i, j = divmod(int(q), n)
        for unit in (1.0, 1j):
            gp = g_t.clone()
            gm = g_t.clone()
            gp[i, j] += unit * h
            gm[i, j] -= unit * h
            fd = dloss(gp, c_t, gm, c_t)
            an = g32[i, j].real if unit == 1.0 else g32[i, j].imag
            errs.append(abs(fd - an) / max(abs(an), 1e-3 * np.abs(g32).max()))
    assert max(errs) < 1e-3, max(errs)
    errs = []
    for p in rng.choice(M, 32, replace=False):
        for ax in range(2):
            cp = c_t.clone()
            cm = c_t.clone()
            cp[p, ax] += h
            cm[p, ax] -= h
            fd = dloss(g_t, cp, g_t, cm)
            an = c32[p, ax]
            errs.append(abs(fd - an) / max(abs(an), 1e-3 * np.abs(c32).max()))
    errs = sorted(errs)
    # one entry may sit next to a kernel piece boundary (as for complex128)
    assert errs[-2] < 1e-3, errs[-5:]


def _c5_request(rng):
    """One request of the mixed distribution: dim 1-3, grid size per axis
    log-uniform in 64..4096 (1-D up to 1e6), M log-uniform in 1e3..1e7, eps
    log-uniform in 1e-10..1e-3, uniform or 8-cluster points, either
    transform type."""
    nd = int(rng.integers(1, 4))
    hi = 1e6 if nd == 1 else 4096
    shape = tuple(int(np.exp(rng.uniform(np.log(64), np.log(hi))))
                  for _ in range(nd))
    M = int(10 ** rng.uniform(3, 7))
    eps = float(10 ** rng.uniform(-10, -3))
    clustered = bool(rng.random() < 0.5)
    nu2u = bool(rng.random() < 0.5)
    return dict(shape=shape, M=M, eps=eps, clustered=clustered, nu2u=nu2u)


def test_c5_learned_choices_valid(gpu):
    """1,000 requests of the configured distribution (3-D grids up to
    4096^3 included: choosing allocates nothing): the learned setting is a
    kernel-table row of the request's dimensionality and precision whose
    tabulated accuracy reaches eps, with oversampling in the request's
    [sigma_min, sigma_max], support 4..16, an even oversampled length of at
    least 16 per axis and a tile the LDS region fits; where the network's
    output is not used, the setting is ducc0's analytic one."""
    import ctypes as ct
    from ducc_amd import _lib, chooser
    if not chooser.available():
        pytest.skip("no trained weights")
    chooser.load()
    db = np.loadtxt(os.path.join(os.path.dirname(chooser.__file__), "data",
                                 "kernel_db.csv"), delimiter=",",
                    comments="#")
    rng = np.random.default_rng(105)
    used = 0
    for q in range(1000):
        r = _c5_request(rng)
        nd = len(r["shape"])
        f64 = r["eps"] < 1e-6
        shp = np.asarray(r["shape"], np.int64)
        res = {}
        for ch in ("learned", "analytic"):
            opt = _lib.options(chooser=ch)
            inf = _lib.DmxPlanInfo()
            _lib.check(_lib.lib().dmx_choose(
                nd, shp.ctypes.data, r["M"], r["eps"], int(f64),
                int(r["nu2u"]), 1.2, 2.0, ct.byref(opt), ct.byref(inf)))
            res[ch] = inf.as_dict()
        f64 = f64 or res["analytic"]["fp64"]
        L = res["learned"]
        k = L["kernel_idx"]
        row = db[k]
        assert int(row[5]) == nd and int(row[6]) == (0 if L["fp64"] else 1)
        # FP64 for tight tolerances, or where the single-precision table
        # cannot reach eps inside the oversampling range
        want64 = f64 or len(O.available_kernels(r["eps"], nd, True, 1.2,
                                                2.0)) == 0
        assert L["fp64"] == want64, r
        assert row[2] <= r["eps"], (r, L)
        assert 1.2 - 1e-9 <= L["ofactor"] <= 2.0 + 1e-9
        assert 4 <= L["W"] <= 16
        assert L["W"] <= (16 if f64 else 8)
        for a in range(nd):
            assert L["nover"][a] % 2 == 0 and L["nover"][a] >= 16
            assert L["nover"][a] >= 2 * ((L["W"] + 1) // 2)
        assert 2 <= L["log2tile"] <= 10
        if L["chooser_used"]:
            used += 1
        else:
            A = res["analytic"]
            assert (k, L["nover"]) == (A["kernel_idx"], A["nover"])
    assert used > 0


def test_c5_learned_accuracy_sample(gpu):
    """A sample of the same distribution (the first 24 requests whose
    oversampled grid stays below 2^27 cells and M below 2e6, to bound the
    run time) transformed with the learned choice (spot check on): every
    result within the reference tests' bound of the FP64 direct sum on 64
    sampled modes (type 1) or points (type 2)."""
    from ducc_amd import chooser, nufft
    if not chooser.available():
        pytest.skip("no trained weights")
    chooser.load()
    rng = np.random.default_rng(1050)
    done = 0
    while done < 24:
        r = _c5_request(rng)
        nd = len(r["shape"])
        if np.prod(r["shape"]) * 2.0 ** nd * 4 > 2 ** 27 or r["M"] > 2e6:
            continue
        M, shape, eps = r["M"], r["shape"], r["eps"]
        if r["clustered"]:
            c = _clustered(rng, M, nd)
        else:
            c = rng.uniform(-np.pi, np.pi, (M, nd))
        dt = np.complex128 if eps < 1e-6 else np.complex64
        ct_ = torch.from_numpy(c).cuda()
        info = {}
        if r["nu2u"]:
            p = _gauss(rng, M, dt)
            out = torch.empty(shape, dtype=torch.from_numpy(p).dtype,
                              device="cuda")
            nufft.nu2u(points=torch.from_numpy(p).cuda(), coord=ct_,
                       forward=True, epsilon=eps, out=out, chooser="learned",
                       info=info)
            sel = rng.choice(int(np.prod(shape)),
                             min(64, int(np.prod(shape))), replace=False)
            ref = nufft.nudft_direct(1, c, p.astype(np.complex128), shape,
                                     True, sel)
            got = out.cpu().numpy().ravel()[sel]
        else:
            g = _gauss(rng, shape, dt)
            out = nufft.u2nu(grid=torch.from_numpy(g).cuda(), coord=ct_,
                             forward=True, epsilon=eps, chooser="learned",
                             info=info)
            sel = rng.choice(M, min(64, M), replace=False)
            ref = nufft.nudft_direct(2, c, g.astype(np.complex128), shape,
                                     True, sel)
            got = out.cpu().numpy()[sel]
        mul = {1: 4, 2: 2, 3: 1}[nd]
        # single precision adds its own rounding floor (the tabulated
        # accuracy of the FP32 kernels stops near 1e-6)
        bound = mul * max(eps, 1e-6 if dt == np.complex64 else 0)
        e = _l2(got, ref)
        assert e <= bound, (r, info, e)
        done += 1
