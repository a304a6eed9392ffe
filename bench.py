#!/usr/bin/env python
"""Throughput of the MI355X NUFFT on the 2-D type-1 + type-2 workload.

Workload: 2048 x 2048 uniform grid, M = 10,000,000 nonuniform points drawn
from 8 Gaussian clusters (sigma 0.3 rad, centres uniform, wrapped to
[-pi, pi)), float64 coordinates, complex64 values, eps = 1e-5.  One step is
one nu2u (type 1) plus one u2nu (type 2) on that geometry, inputs resident
in HBM.  The headline value is 2*M*steps / time for the fast path (FP32
compute with FP64 folding, learned parameter choice when a trained network
is present); the plain FP64 path (complex128, ducc0's analytic parameters)
is timed on the same geometry for the speed-up, and ducc0 itself (built from
source into oracle/_ref) is timed on the host cores as the CPU baseline.

Prints one JSON line.  Usage: python bench.py [--gpus 1] [--steps K]
[--warmup W]; `--workload c2c` measures the ducc0.fft.c2c drop-in instead
(see bench_c2c).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM_GBS = 8000.0
# ds_add_f64 with the lanes of each quarter-wave on distinct bank classes:
# a rate measured on MI355X with a microbenchmark (profiles/lds_atomics.md),
# not a published figure; random addresses reach a third of it
PEAK_LDS_F64_ATOMICS = 4.66e12
# ds_add_u64 on distinct cells, the same microbenchmark: the FP32 path's
# fixed-point spreading (eps >= 1e-5) adds one per tap
PEAK_LDS_U64_ATOMICS = 5.06e12
FIX_MIN_EPS = 1e-5


def clustered_points(rng, M, d=2, ncl=8, sig=0.3):
    cen = rng.uniform(-np.pi, np.pi, (ncl, d))
    lab = rng.integers(0, ncl, M)
    x = cen[lab] + rng.normal(0.0, sig, (M, d))
    return (x + np.pi) % (2 * np.pi) - np.pi


def complex_gauss(rng, shape, dtype):
    z = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    return (z * np.sqrt(0.5)).astype(dtype)


def alg_bytes(phase, info, M, d, sv):
    """Algorithmic HBM bytes of one launch of a single-kernel phase:
    spreading / interpolation read each point's coordinates (8 B per axis),
    its sort index (4 B) and value (sv B) once and write / read the
    oversampled grid once; an FFT axis pass reads and writes the lines it
    transforms once."""
    nover = info["nover"]
    nuni = info["nuni"]
    n_over = int(np.prod(nover))
    if phase in ("spreading", "interpolation"):
        return M * (d * 8 + 4 + sv) + n_over * sv
    if phase.startswith("FFT axis"):
        return None
    return None


def fft_pass_bytes(info, type1, sv):
    """Bytes of each FFT axis pass with the reference's pruning: the lines
    along axis k run over the kept bands (nuni of nover) of every later
    axis and over all of every earlier axis, for both transform types."""
    nover, nuni = info["nover"], info["nuni"]
    nd = len(nover)
    out = {}
    for k in range(nd):
        n = 1
        for a in range(nd):
            n *= nuni[a] if a > k else nover[a]
        out[f"FFT axis {k}"] = 2 * n * sv
    return out


def l2rel(a, b):
    """ducc0's l2error: sqrt(sum|a-b|^2 / max(sum|a|^2, sum|b|^2))"""
    a = np.asarray(a, np.complex128).ravel()
    b = np.asarray(b, np.complex128).ravel()
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) /
                         max(np.sum(np.abs(a) ** 2), np.sum(np.abs(b) ** 2))))


def time_steps(step, steps, warmup, sync):
    for _ in range(warmup):
        step()
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    sync()
    return time.perf_counter() - t0


def _median_time(fn, reps=5):
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def cpu_baseline(coords, pts, grid, eps, M_total, fast=None):
    """ducc0 (compiled from the reference sources into oracle/_ref) on all
    host cores, on the same workload: after one warm-up, the median of 5
    unplanned nu2u + u2nu passes (the value), and the median of 5 planned
    execute-only passes of ducc0's plan (its plan needs coordinates in the
    values' precision: float32 here).  With `fast` = (type-1 result, type-2
    result) of the GPU fast path, also their L2 errors against ducc0's own
    double-precision nu2u / u2nu of the same inputs."""
    try:
        from oracle import ref as R
    except Exception:
        return None
    if not R.available():
        return None
    nth = os.cpu_count() or 1
    shape = grid.shape
    R.nu2u(pts, coords, True, eps, shape, nthreads=nth)
    R.u2nu(grid, coords, False, eps, nthreads=nth)
    t1 = _median_time(lambda: R.nu2u(pts, coords, True, eps, shape,
                                     nthreads=nth))
    t2 = _median_time(lambda: R.u2nu(grid, coords, False, eps, nthreads=nth))
    planned = None
    try:
        c32 = coords.astype(np.float32)
        t0 = time.perf_counter()
        p1 = R.Plan(True, c32, shape, eps, True, nthreads=nth,
                    sigma_min=1.2, sigma_max=2.51)
        p2 = R.Plan(False, c32, shape, eps, True, nthreads=nth,
                    sigma_min=1.2, sigma_max=2.51)
        tp = time.perf_counter() - t0
        p1.nu2u(pts)
        p2.u2nu(grid, False)
        e1 = _median_time(lambda: p1.nu2u(pts))
        e2 = _median_time(lambda: p2.u2nu(grid, False))
        planned = dict(value=2 * len(pts) / (e1 + e2), unit="points/s",
                       plan_s=round(tp, 4), nu2u_s=round(e1, 4),
                       u2nu_s=round(e2, 4))
        del p1, p2
    except Exception as e:  # the baseline must not sink the bench line
        planned = dict(error=str(e))
    err = None
    if fast is not None:
        # eps 1e-5 in double precision (W from ducc0's FP64 table)
        r1 = R.nu2u(pts.astype(np.complex128), coords, True, eps, shape,
                    nthreads=nth)
        r2 = R.u2nu(grid.astype(np.complex128), coords, False, eps,
                    nthreads=nth)
        err = dict(nu2u=l2rel(fast[0], r1), u2nu=l2rel(fast[1], r2))
    return dict(value=2 * len(pts) / (t1 + t2), unit="points/s", cores=nth,
                kind="reference", err_fast_vs_ducc0_fp64=err,
                planned_execute_only=planned,
                sample=(f"ducc0 nu2u + u2nu, full workload (M={len(pts)}, "
                        f"{shape[0]}x{shape[1]}, eps={eps}, complex64, "
                        f"{nth} threads): median of 5 unplanned calls each "
                        f"after a warm-up, nu2u {t1:.3f} s, u2nu {t2:.3f} s; "
                        f"planned execute-only in planned_execute_only"))


def c2c_cpu(R, a, nthreads, reps):
    """ducc0::c2c (oracle/_ref) on the host: mean seconds of `reps` calls"""
    R.c2c(a, nthreads=nthreads)
    t0 = time.perf_counter()
    for _ in range(reps):
        out = R.c2c(a, nthreads=nthreads)
    return (time.perf_counter() - t0) / reps, out


def fibonacci_lattice(npoints):
    """(theta, phi) of a Fibonacci lattice (python/demos/sht_general.py:20-26)"""
    gold = (1. + np.sqrt(5.)) / 2.
    r = np.arange(npoints, dtype=np.float64)
    loc = np.empty((npoints, 2))
    loc[:, 0] = np.arccos(1. - (2 / npoints * r))
    loc[:, 1] = np.mod(r / gold, 1.) * (2 * np.pi)
    return loc


def random_alm(lmax, rng):
    """random a_lm of lmax = mmax in the default layout (m-major, m = 0 first),
    real and imaginary parts uniform in [-1, 1), the m = 0 entries real
    (python/test/test_sht.py:31-40)"""
    n = (lmax + 1) * (lmax + 2) // 2
    res = rng.uniform(-1., 1., n) + 1j * rng.uniform(-1., 1., n)
    res[:lmax + 1].imag = 0.
    return res


def bench_sht(args):
    """ducc0.sht.synthesis_general + adjoint_synthesis_general on the
    reference demo's sky (python/demos/sht_general.py:29-41: lmax = mmax =
    1187, 2,000,000 points of a Fibonacci lattice, random a_lm), complex128
    a_lm, eps 1e-7; spin 0 by default, --spin S for a spin-S field (E and B,
    mode STANDARD).  One step is one synthesis plus one adjoint with inputs
    on the device; ducc0 itself (oracle/_ref) on all host cores is the CPU
    baseline (median of 3 after a warm-up)."""
    import torch
    from ducc_amd import _lib, sht
    dev = torch.device("cuda:0")
    lmax, npt, eps, spin = 1187, 2_000_000, 1e-7, args.spin
    nc = 1 if spin == 0 else 2
    rng = np.random.default_rng(42)
    a_np = np.stack([random_alm(lmax, rng) for _ in range(nc)])
    ofs = 0
    for s_ in range(spin):  # l < spin: zero (test_sht.py:31-40)
        a_np[:, ofs:ofs + spin - s_] = 0
        ofs += lmax + 1 - s_
    loc_np = fibonacci_lattice(npt)
    r_np = rng.uniform(-1, 1, (nc, npt))
    a = torch.from_numpy(a_np).to(dev)
    loc = torch.from_numpy(loc_np).to(dev)
    r = torch.from_numpy(r_np).to(dev)
    mp = torch.empty((nc, npt), dtype=torch.float64, device=dev)
    at = torch.empty_like(a)
    sync = torch.cuda.synchronize

    def step():
        sht.synthesis_general(alm=a, spin=spin, lmax=lmax, loc=loc,
                              epsilon=eps, map=mp)
        sht.adjoint_synthesis_general(map=r, spin=spin, lmax=lmax, loc=loc,
                                      epsilon=eps, alm=at)
    steps = max(1, args.steps // 2)
    t = time_steps(step, steps, max(1, args.warmup // 2), sync)
    _lib.profile_enable(True)
    _lib.profile_reset()
    step()
    sync()
    prof = _lib.profile_read()
    _lib.profile_enable(False)
    cpu = None
    if not args.no_cpu:
        try:
            from oracle import ref as R
            if R.available():
                nth = os.cpu_count() or 1
                na = a_np.shape[1]
                f_ref = R.sht_synthesis_general(a_np, lmax, loc_np, eps,
                                                nthreads=nth, spin=spin)
                b_ref = R.sht_adjoint_synthesis_general(
                    r_np, lmax, loc_np, eps, na, nthreads=nth, spin=spin)
                ts = _median_time(lambda: R.sht_synthesis_general(
                    a_np, lmax, loc_np, eps, nthreads=nth, spin=spin), reps=3)
                ta = _median_time(lambda: R.sht_adjoint_synthesis_general(
                    r_np, lmax, loc_np, eps, na, nthreads=nth, spin=spin),
                    reps=3)
                cpu = dict(value=npt * 2 / (ts + ta), unit="points/s",
                           cores=nth, kind="reference",
                           synthesis_s=round(ts, 4), adjoint_s=round(ta, 4),
                           err_vs_ducc0=dict(
                               synthesis=l2rel(mp.cpu().numpy().ravel(),
                                               np.ravel(f_ref)),
                               adjoint=l2rel(at.cpu().numpy().ravel(),
                                             np.ravel(b_ref))),
                           sample="ducc0 synthesis_general + "
                                  "adjoint_synthesis_general, same inputs, "
                                  "median of 3 after a warm-up")
        except Exception as e:
            cpu = dict(error=str(e))
    line = dict(
        metric="points/s, spherical harmonic synthesis + adjoint at "
               f"arbitrary points (spin {spin})",
        value=round(2 * npt * steps / t, 1), unit="points/s", n_gpus=1,
        steps=steps, warmup=max(1, args.warmup // 2),
        ms_per_step=round(1e3 * t / steps, 4), higher_is_better=True,
        scaling="weak", vs_baseline=None, dtype="fp64", data="synthetic",
        config=dict(workload="ducc0.sht.synthesis_general + "
                             f"adjoint_synthesis_general, spin {spin}, lmax "
                             "= mmax = 1187, 2,000,000 Fibonacci-lattice "
                             "points, eps 1e-7, complex128 a_lm"),
        cpu_baseline=cpu,
        phases_ms={k: round(v[0], 4) for k, v in prof.items()})
    print(json.dumps(line))


def bench_sht_pseudo(args):
    """ducc0.sht.pseudo_analysis_general as the reference demo runs it
    (python/demos/sht_general.py:29-45): a random spin-2 sky at lmax = mmax
    = 1187 synthesized at eps 1e-12 on 2,000,000 points of a Fibonacci
    lattice, then analysed at eps 1e-7 with at most 20 LSMR iterations.
    One step is one pseudo_analysis_general with the map and the points on
    the device; ducc0 itself (oracle/_ref) on all host cores, on the same
    map, is the CPU baseline (one call after a warm-up: about 2 s)."""
    import torch
    from ducc_amd import _lib, sht
    dev = torch.device("cuda:0")
    lmax, npt, eps, spin, maxiter = 1187, 2_000_000, 1e-7, 2, 20
    rng = np.random.default_rng(42)
    a_np = np.stack([random_alm(lmax, rng) for _ in range(2)])
    ofs = 0
    for s_ in range(spin):  # l < spin: zero (sht_general.py:8-17)
        a_np[:, ofs:ofs + spin - s_] = 0
        ofs += lmax + 1 - s_
    loc_np = fibonacci_lattice(npt)
    a = torch.from_numpy(a_np).to(dev)
    loc = torch.from_numpy(loc_np).to(dev)
    mp = torch.empty((2, npt), dtype=torch.float64, device=dev)
    sht.synthesis_general(alm=a, spin=spin, lmax=lmax, loc=loc,
                          epsilon=1e-12, map=mp)
    at = torch.empty_like(a)
    res = {}
    sync = torch.cuda.synchronize

    def step():
        res["r"] = sht.pseudo_analysis_general(
            map=mp, loc=loc, lmax=lmax, spin=spin, nthreads=1, epsilon=eps,
            maxiter=maxiter, alm=at)
    steps = max(1, min(args.steps, 3))
    warm = 1
    t = time_steps(step, steps, warm, sync)
    _, istop, itn, rnorm, _ = res["r"]
    err_truth = l2rel(at.cpu().numpy(), a_np)
    cpu = None
    if not args.no_cpu:
        try:
            from oracle import ref as R
            if R.available():
                nth = os.cpu_count() or 1
                m_np = mp.cpu().numpy()
                R.sht_pseudo_analysis_general(m_np, lmax, loc_np, spin, 2,
                                              eps, a_np.shape[1],
                                              nthreads=nth)
                t0 = time.perf_counter()
                ref = R.sht_pseudo_analysis_general(m_np, lmax, loc_np, spin,
                                                    maxiter, eps,
                                                    a_np.shape[1],
                                                    nthreads=nth)
                tc = time.perf_counter() - t0
                cpu = dict(value=1. / tc, unit="analyses/s", cores=nth,
                           kind="reference", seconds=round(tc, 3),
                           istop=int(ref[1]), itn=int(ref[2]),
                           rnorm=float(ref[3]),
                           err_vs_ducc0=l2rel(at.cpu().numpy(), ref[0]),
                           sample="ducc0 pseudo_analysis_general, same map "
                                  "and points, one call after a 2-iteration "
                                  "warm-up")
        except Exception as e:
            cpu = dict(error=str(e))
    line = dict(
        metric="analyses/s, spherical harmonic pseudo-analysis at arbitrary "
               "points (spin 2, LSMR)",
        value=round(steps / t, 3), unit="analyses/s", n_gpus=1,
        steps=steps, warmup=warm, ms_per_step=round(1e3 * t / steps, 3),
        higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="fp64", data="synthetic",
        config=dict(workload="ducc0.sht.pseudo_analysis_general, spin 2, "
                             "lmax = mmax = 1187, 2,000,000 Fibonacci-lattice "
                             "points, eps 1e-7, maxiter 20"),
        istop=int(istop), itn=int(itn), rnorm=float(rnorm),
        err_vs_input_alm=err_truth, cpu_baseline=cpu)
    print(json.dumps(line))


def bench_wgridder(args):
    """ducc0.wgridder vis2dirty + dirty2vis with w-gridding on a synthetic
    observation shaped like the reference's benchmark
    (python/demos/wgridder_bench.py: eps 1e-4, single precision, w-gridding,
    image size from the field of view and the longest baseline, :11-17):
    20,000 rows x 64 channels from 1 to 2 GHz (1.28e6 visibilities), u and v
    uniform in +-1500 m, w in +-500 m, a 5 degree field, complex64.  One
    step is one gridding plus one degridding with inputs on the device;
    ducc0 itself (oracle/_ref) on all host cores is the CPU baseline
    (median of 3 after a warm-up)."""
    import torch
    from ducc_amd import _lib, wgridder as W
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(7)
    nrow, nchan, eps, fov = 20_000, 64, 1e-4, 5.0
    c = 299792458.
    freq = 1e9 + np.arange(nchan) * (1e9 / nchan)
    uvw = (rng.random((nrow, 3)) - 0.5) * np.array([3000., 3000., 1000.])
    vis = ((rng.random((nrow, nchan)) - 0.5) +
           1j * (rng.random((nrow, nchan)) - 0.5)).astype(np.complex64)
    wgt = rng.uniform(0.5, 1.5, (nrow, nchan)).astype(np.float32)
    bl = np.sqrt(np.sum(uvw ** 2, axis=1))
    maxbl = np.max(bl[:, None] * freq[None, :] / c)
    npix = int(2 * fov * np.pi / 180 * maxbl) + 1
    npix += npix % 2
    pix = fov / npix * np.pi / 180
    dirty = (rng.random((npix, npix)) - 0.5).astype(np.float32)
    t = lambda x: torch.from_numpy(x).to(dev)
    uvw_d, freq_d, vis_d, wgt_d, dirty_d = t(uvw), t(freq), t(vis), t(wgt), \
        t(dirty)
    out_d = torch.empty((npix, npix), dtype=torch.float32, device=dev)
    vout_d = torch.empty((nrow, nchan), dtype=torch.complex64, device=dev)
    sync = torch.cuda.synchronize
    kw = dict(pixsize_x=pix, pixsize_y=pix, epsilon=eps, do_wgridding=True)

    def step():
        W.vis2dirty(uvw=uvw_d, freq=freq_d, vis=vis_d, wgt=wgt_d,
                    dirty=out_d, **kw)
        W.dirty2vis(uvw=uvw_d, freq=freq_d, dirty=dirty_d, wgt=wgt_d,
                    vis=vout_d, **kw)
    steps = max(1, args.steps // 2)
    tt = time_steps(step, steps, max(1, args.warmup // 2), sync)
    _lib.profile_enable(True)
    _lib.profile_reset()
    step()
    sync()
    prof = _lib.profile_read()
    _lib.profile_enable(False)
    nvis = nrow * nchan
    cpu = None
    if not args.no_cpu:
        try:
            from oracle import ref as R
            if R.available():
                nth = os.cpu_count() or 1
                d_ref = R.wgridder(False, uvw, freq, vis, wgt, None, npix,
                                   npix, pix, pix, eps, True, nthreads=nth)
                v_ref = R.wgridder(True, uvw, freq, dirty, wgt, None, npix,
                                   npix, pix, pix, eps, True, nthreads=nth)
                tg = _median_time(lambda: R.wgridder(
                    False, uvw, freq, vis, wgt, None, npix, npix, pix, pix,
                    eps, True, nthreads=nth), reps=3)
                td = _median_time(lambda: R.wgridder(
                    True, uvw, freq, dirty, wgt, None, npix, npix, pix, pix,
                    eps, True, nthreads=nth), reps=3)
                cpu = dict(value=2 * nvis / (tg + td), unit="visibilities/s",
                           cores=nth, kind="reference",
                           vis2dirty_s=round(tg, 4),
                           dirty2vis_s=round(td, 4),
                           err_vs_ducc0=dict(
                               vis2dirty=l2rel(out_d.cpu().numpy(), d_ref),
                               dirty2vis=l2rel(vout_d.cpu().numpy(), v_ref)),
                           sample="ducc0 vis2dirty + dirty2vis (ms2dirty / "
                                  "dirty2ms), same inputs, median of 3 after "
                                  "a warm-up")
        except Exception as e:
            cpu = dict(error=str(e))
    line = dict(
        metric="visibilities/s, w-gridding vis2dirty + dirty2vis",
        value=round(2 * nvis * steps / tt, 1), unit="visibilities/s",
        n_gpus=1, steps=steps, warmup=max(1, args.warmup // 2),
        ms_per_step=round(1e3 * tt / steps, 4), higher_is_better=True,
        scaling="weak", vs_baseline=None, dtype="fp32", data="synthetic",
        config=dict(workload=f"ducc0.wgridder vis2dirty + dirty2vis, "
                             f"w-gridding, {nrow} rows x {nchan} channels, "
                             f"{npix}x{npix} pixels, 5 deg field, eps 1e-4, "
                             f"complex64"),
        cpu_baseline=cpu,
        phases_ms={k: round(v[0], 4) for k, v in prof.items()})
    print(json.dumps(line))


def bench_c2c(args):
    """ducc0.fft.c2c drop-in: in-place forward c2c of a 4096 x 4096
    complex64 array (the headline of this mode, HBM-bound: each transformed
    axis reads and writes the array once, 2 * N * 8 bytes per axis), plus the
    reference's own FFT benchmark shapes (python/demos/fft_bench.py:127-170:
    random good_size extents in [nmax/3, nmax], nmax 8192 / 2048 / 256 for
    1-D / 2-D / 3-D, complex128 and complex64, in place), each timed on the
    GPU and with ducc0 itself on one host thread (as fft_bench.py runs it)."""
    import torch
    from ducc_amd import _lib, fft
    try:
        from oracle import ref as R
        have_ref = R.available() and not args.no_cpu
    except Exception:
        have_ref = False
    dev = torch.device("cuda:0")
    sync = torch.cuda.synchronize
    shp = (4096, 4096)
    N = shp[0] * shp[1]
    g = torch.Generator(device="cuda").manual_seed(21)
    x = torch.randn(shp, generator=g, dtype=torch.complex64, device=dev)
    work = x.clone()

    def step():
        fft.c2c(work, out=work, forward=True)
    t = time_steps(step, args.steps, args.warmup, sync)
    ms = 1e3 * t / args.steps
    _lib.profile_enable(True)
    _lib.profile_reset()
    for _ in range(args.steps):
        step()
    sync()
    prof = _lib.profile_read()
    _lib.profile_enable(False)
    per = {k: v[0] / v[1] for k, v in prof.items()}
    alg = 2 * N * 8  # per axis: read + write of the array
    axis_ms = {k: v for k, v in per.items() if "FFT axis" in k}
    dom = max(axis_ms, key=axis_ms.get)
    ach = alg / (axis_ms[dom] * 1e-3) / 1e9
    value = 2 * alg / (ms * 1e-3) / 1e9
    # HBM bytes of the phase from the profiler's FETCH_SIZE + WRITE_SIZE of
    # this transform (calibrated by a plain copy)
    traffic = None
    tp = os.path.join(HERE, "profiles", "c2c_traffic.json")
    if os.path.exists(tp):
        try:
            traffic = json.load(open(tp)).get(dom)
        except Exception:
            traffic = None
    roofline = dict(bound="hbm", kernel=dom, achieved=round(ach, 1),
                    peak=PEAK_HBM_GBS, unit="GB/s",
                    frac=round(ach / PEAK_HBM_GBS, 4), traffic=traffic,
                    alg_bytes_per_launch=alg,
                    avg_launch_ms=round(axis_ms[dom], 4))
    if traffic:
        hb = traffic / (axis_ms[dom] * 1e-3) / 1e9
        roofline["traffic_GBps"] = round(hb, 1)
        roofline["traffic_frac"] = round(hb / PEAK_HBM_GBS, 4)
    # accuracy against numpy's FP64 FFT of the same input
    xin = x.cpu().numpy()
    work.copy_(x)
    step()
    got = work.cpu().numpy()
    ref64 = np.fft.fft2(xin.astype(np.complex128))
    err = l2rel(got, ref64)
    cpu = None
    if have_ref:
        nth = os.cpu_count() or 1
        tc, rc = c2c_cpu(R, xin, nth, 3)
        cpu = dict(value=round(2 * alg / tc / 1e9, 2), unit="GB/s",
                   cores=nth, kind="reference",
                   err_gpu_vs_ducc0=l2rel(got, rc),
                   sample=f"ducc0 c2c of the same 4096x4096 complex64 array, "
                          f"{nth} threads, mean of 3 calls after a warm-up: "
                          f"{1e3 * tc:.2f} ms")
    # the reference's benchmark shapes
    rng = np.random.default_rng(17)
    rows = []
    for ndim, nmax in ((1, 8192), (2, 2048), (3, 256)):
        for tp in (np.complex128, np.complex64):
            for _ in range(3):
                sh = tuple(fft.good_size(int(v)) for v in
                           rng.integers(nmax // 3, nmax + 1, ndim))
                a = (rng.random(sh) - 0.5 + 1j * (rng.random(sh) - 0.5)) \
                    .astype(tp)
                at = torch.from_numpy(a).to(dev)
                w = at.clone()

                def st():
                    fft.c2c(w, out=w, forward=True)
                tg = time_steps(st, 10, 2, sync) / 10
                w.copy_(at)
                st()
                row = dict(shape=list(sh), dtype=np.dtype(tp).name,
                           gpu_ms=round(1e3 * tg, 4))
                if have_ref:
                    tcpu, rr = c2c_cpu(R, a, 1, 3)
                    row.update(ducc0_1thread_ms=round(1e3 * tcpu, 4),
                               speedup=round(tcpu / tg, 1),
                               l2err_vs_ducc0=l2rel(w.cpu().numpy(), rr))
                rows.append(row)
    line = dict(
        metric="c2c GB/s (algorithmic: read + write of the array per "
               "transformed axis), 4096x4096 complex64, forward, in place",
        value=round(value, 1), unit="GB/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=round(ms, 4), higher_is_better=True,
        scaling="weak", vs_baseline=None, dtype="fp32", data="synthetic",
        config=dict(workload="ducc0.fft.c2c, 2-D 4096x4096 complex64, "
                             "forward, in place on the device"),
        roofline=roofline,
        cpu_baseline=cpu, err_vs_numpy_fp64=err,
        phases_ms={k: round(v, 4) for k, v in per.items()},
        fft_bench_shapes=rows)
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="nufft",
                    choices=["nufft", "c2c", "sht", "wgridder"],
                    help="nufft: the 2-D type-1 + type-2 workload (default); "
                         "c2c: the ducc0.fft.c2c drop-in; sht: spherical "
                         "harmonic synthesis at arbitrary points; wgridder: "
                         "radio-interferometric gridding + degridding")
    ap.add_argument("--spin", type=int, default=0,
                    help="sht workload: spin of the field (0, or > 0 for "
                         "E and B)")
    ap.add_argument("--pseudo", action="store_true",
                    help="sht workload: pseudo_analysis_general as the "
                         "reference demo runs it (spin 2, 20 iterations)")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--M", type=int, default=10_000_000)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--eps", type=float, default=1e-5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--chooser", default="auto",
                    choices=["auto", "learned", "analytic"])
    ap.add_argument("--point-records", action="store_true",
                    help="nufft workload: sort point records instead of a "
                         "permutation (DMX_POINT_RECORDS; off by default)")
    ap.add_argument("--calib", action="store_true",
                    help="run one 256 MiB device copy first (calibrates HBM "
                         "byte counters of a profiler run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path "
              "(the NUFFT would shard points by tile ranges; not built); run "
              "with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    dev = torch.device("cuda:0")
    if args.calib:
        # a plain copy of known size: 256 MiB read + 256 MiB written
        x = torch.empty(1 << 26, dtype=torch.float32, device=dev).fill_(1.0)
        y = x.clone()
        torch.cuda.synchronize()
        del x, y
    if args.workload == "c2c":
        bench_c2c(args)
        return
    if args.workload == "sht":
        (bench_sht_pseudo if args.pseudo else bench_sht)(args)
        return
    if args.workload == "wgridder":
        bench_wgridder(args)
        return
    from ducc_amd import _lib, nufft
    from ducc_amd import chooser as learned
    rng = np.random.default_rng(102)
    M, n, eps = args.M, args.n, args.eps
    c_np = clustered_points(rng, M)
    p_np = complex_gauss(rng, M, np.complex64)
    g_np = complex_gauss(rng, (n, n), np.complex64)
    c = torch.from_numpy(c_np).to(dev)
    p32 = torch.from_numpy(p_np).to(dev)
    g32 = torch.from_numpy(g_np).to(dev)
    og32 = torch.empty((n, n), dtype=torch.complex64, device=dev)
    op32 = torch.empty(M, dtype=torch.complex64, device=dev)
    chooser = args.chooser
    if chooser == "auto":
        chooser = "learned" if learned.available() else "analytic"
    sync = torch.cuda.synchronize

    info1, info2 = {}, {}
    # --point-records: the opt-in record sort instead of the index sort
    rec = dict(point_records=True) if args.point_records else {}

    def step_fast():
        nufft.nu2u(points=p32, coord=c, forward=True, epsilon=eps, out=og32,
                   chooser=chooser, info=info1, **rec)
        nufft.u2nu(grid=g32, coord=c, forward=False, epsilon=eps, out=op32,
                   chooser=chooser, info=info2, **rec)

    # timed with the phase profiler off; the phase times come from a second
    # pass of the same steps with HIP events around each phase
    t = time_steps(step_fast, args.steps, args.warmup, sync)
    value = 2 * M * args.steps / t
    _lib.profile_enable(True)
    _lib.profile_reset()
    for _ in range(args.steps):
        step_fast()
    sync()
    prof = _lib.profile_read()
    _lib.profile_enable(False)

    # planned execute-only (sort once per plan, as ducc0.nufft.plan), the
    # same two transforms per step; a plan takes coordinates in the values'
    # precision (nufft_pymod.cc:145-150), float32 here
    c32 = c.to(torch.float32)
    pl1 = nufft.plan(nu2u=True, coord=c32, grid_shape=(n, n), epsilon=eps,
                     sigma_min=1.2, sigma_max=2.51, chooser=chooser)
    pl2 = nufft.plan(nu2u=False, coord=c32, grid_shape=(n, n), epsilon=eps,
                     sigma_min=1.2, sigma_max=2.51, chooser=chooser)

    # own output buffers: og32 / op32 keep the unplanned results, whose
    # accuracy is checked below
    ogp = torch.empty_like(og32)
    opp = torch.empty_like(op32)

    def step_planned():
        pl1.nu2u(points=p32, forward=True, out=ogp)
        pl2.u2nu(grid=g32, forward=False, out=opp)
    tpl = time_steps(step_planned, args.steps, args.warmup, sync)
    planned = dict(value=2 * M * args.steps / tpl,
                   ms_per_step=1e3 * tpl / args.steps,
                   note="ducc0.nufft.plan semantics: index built once per "
                        "plan (float32 coordinates); each step executes "
                        "nu2u + u2nu")
    del pl1, pl2, c32, ogp, opp

    # plain FP64 path on the same geometry
    fp64 = None
    if not args.no_fp64:
        p64 = p32.to(torch.complex128)
        g64 = g32.to(torch.complex128)
        og64 = torch.empty((n, n), dtype=torch.complex128, device=dev)
        op64 = torch.empty(M, dtype=torch.complex128, device=dev)
        info3, info4 = {}, {}

        def step_fp64():
            nufft.nu2u(points=p64, coord=c, forward=True, epsilon=eps,
                       out=og64, precision="fp64", chooser="analytic",
                       info=info3)
            nufft.u2nu(grid=g64, coord=c, forward=False, epsilon=eps,
                       out=op64, precision="fp64", chooser="analytic",
                       info=info4)
        t64 = time_steps(step_fp64, max(1, args.steps // 2), 1, sync)
        v64 = 2 * M * max(1, args.steps // 2) / t64
        # accuracy of the fast path against the FP64 path and the direct sum
        a1 = og32.cpu().numpy().astype(np.complex128)
        b1 = og64.cpu().numpy()
        a2 = op32.cpu().numpy().astype(np.complex128)
        b2 = op64.cpu().numpy()

        l2 = l2rel
        sel_m = rng.choice(n * n, 256, replace=False)
        sel_p = rng.choice(M, 256, replace=False)
        d1 = nufft.nudft_direct(1, c_np, p_np.astype(np.complex128), (n, n),
                                True, sel_m)
        d2 = nufft.nudft_direct(2, c_np, g_np.astype(np.complex128), (n, n),
                                False, sel_p)
        fp64 = dict(value=v64, ms_per_step=1e3 * t64 / max(1, args.steps // 2),
                    W=info3.get("W"), nover=info3.get("nover"),
                    speedup=value / v64,
                    alg_bytes_per_step=None, bytes_ratio_fp64_over_fp32=None,
                    err_fast_vs_fp64=dict(nu2u=l2(a1, b1), u2nu=l2(a2, b2)),
                    err_vs_direct_nudft=dict(
                        nu2u_fast=l2(a1.ravel()[sel_m], d1),
                        nu2u_fp64=l2(b1.ravel()[sel_m], d1),
                        u2nu_fast=l2(a2[sel_p], d2),
                        u2nu_fp64=l2(b2[sel_p], d2)))
        del p64, g64, og64, op64

    # roofline of the dominant single-kernel phase
    sv = 8
    d = 2
    per_launch = {}
    for ph, (ms, cnt) in prof.items():
        per_launch[ph] = ms / max(cnt, 1)
    cand = {ph: ms for ph, (ms, cnt) in prof.items()
            if ph in ("nu2u/spreading", "u2nu/interpolation")
            or "/FFT axis" in ph}
    dom = max(cand, key=cand.get)
    kind, name = dom.split("/", 1)
    inf = info1 if kind == "nu2u" else info2
    if name.startswith("FFT axis"):
        b = fft_pass_bytes(inf, kind == "nu2u", sv)[name]
    else:
        b = alg_bytes(name, inf, M, d, sv)
    achieved = b / (per_launch[dom] * 1e-3) / 1e9
    # HBM bytes per launch from the profiler's FETCH_SIZE + WRITE_SIZE of
    # this workload (calibrated by a plain copy), when recorded for the
    # parameters in use
    traffic = None
    tp = os.path.join(HERE, "profiles", "traffic.json")
    if os.path.exists(tp) and inf.get("W") == 7 and \
            list(inf.get("nover", [])) == [3388, 3388]:
        try:
            traffic = json.load(open(tp)).get(dom)
        except Exception:
            traffic = None
    roofline = dict(bound="hbm", kernel=dom, achieved=round(achieved, 1),
                    peak=PEAK_HBM_GBS, unit="GB/s",
                    frac=round(achieved / PEAK_HBM_GBS, 4), traffic=traffic,
                    alg_bytes_per_launch=int(b),
                    avg_launch_ms=round(per_launch[dom], 4))
    if traffic:
        hb = traffic / (per_launch[dom] * 1e-3) / 1e9
        roofline["traffic_GBps"] = round(hb, 1)
        roofline["traffic_frac"] = round(hb / PEAK_HBM_GBS, 4)
    if name == "spreading":
        # LDS atomics per point: the FP32 path at eps >= 1e-5 adds each tap
        # as one packed fixed-point ds_add_u64, otherwise as two ds_add_f64
        # (re, im); one ranking atomic per point besides
        fix = not inf.get("fp64") and eps >= FIX_MIN_EPS
        na = M * ((1 if fix else 2) * inf["W"] ** d + 1)
        ra = na / (per_launch[dom] * 1e-3)
        pk = PEAK_LDS_U64_ATOMICS if fix else PEAK_LDS_F64_ATOMICS
        roofline["lds_atomics"] = dict(
            kind="ds_add_u64 (fixed point)" if fix else "ds_add_f64",
            per_launch=int(na), achieved=float(f"{ra:.4g}"),
            peak=pk, unit="atomics/s", frac=round(ra / pk, 4),
            peak_source="self-measured microbenchmark rate "
                        "(profiles/lds_atomics.md), not a published figure")

    # whole-step roofline: per call M*(d*s_c + s_v) + N*s_v +
    # (2d+2)*N_over*s_v (coordinates and values once; the uniform array once;
    # the oversampled grid written, read + written per FFT axis, read once)
    def call_bytes(inf):
        n_over = int(np.prod(inf["nover"]))
        return M * (d * 8 + sv) + n * n * sv + (2 * d + 2) * n_over * sv
    step_b = call_bytes(info1) + call_bytes(info2)
    if fp64 is not None:
        # the same formula with 16-byte values on the FP64 path's grids:
        # an HBM-bound FP32 path can be at most this ratio faster
        def call_bytes64(inf):
            n_over = int(np.prod(inf["nover"]))
            return M * (d * 8 + 16) + n * n * 16 + (2 * d + 2) * n_over * 16
        b64 = call_bytes64(info3) + call_bytes64(info4)
        fp64["alg_bytes_per_step"] = int(b64)
        fp64["bytes_ratio_fp64_over_fp32"] = round(b64 / step_b, 3)
    step_gbs = step_b * args.steps / t / 1e9
    step_roofline = dict(alg_bytes_per_step=int(step_b),
                         achieved=round(step_gbs, 1), unit="GB/s",
                         frac=round(step_gbs / PEAK_HBM_GBS, 4))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(c_np, p_np, g_np, eps, M,
                           fast=(og32.cpu().numpy(), op32.cpu().numpy()))

    line = dict(
        metric="nonuniform pts/s (type1+type2) @eps 1e-5: FP32+learned vs "
               "plain FP64 GPU",
        value=round(value, 1), unit="points/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=round(1e3 * t / args.steps, 4),
        higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="fp32 (fp64 coordinate fold)", data="synthetic",
        config=dict(workload="2D NUFFT type1+type2, 2048x2048 grid, "
                             "M=10,000,000 points in 8 Gaussian clusters "
                             "(sigma 0.3 rad), eps=1e-5, complex64 values, "
                             "float64 coords",
                    M=M, grid=[n, n], eps=eps, chooser=chooser,
                    sort="point records" if args.point_records else "index",
                    nu2u=dict(W=info1.get("W"), nover=info1.get("nover"),
                              ofactor=info1.get("ofactor"),
                              log2tile=info1.get("log2tile"),
                              learned=info1.get("chooser_used")),
                    u2nu=dict(W=info2.get("W"), nover=info2.get("nover"),
                              ofactor=info2.get("ofactor"),
                              log2tile=info2.get("log2tile"),
                              learned=info2.get("chooser_used"))),
        roofline=roofline, step_roofline=step_roofline, cpu_baseline=cpu,
        fp64_plain=fp64, planned=planned,
        phases_ms_per_step={k: round(v[0] / args.steps, 4)
                            for k, v in prof.items()})
    print(json.dumps(line))


if __name__ == "__main__":
    main()
