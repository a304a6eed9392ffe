"""Measurements of the secondary configurations on one GPU (the headline
2-D type-1+type-2 workload is bench.py):

  c1  2-D type 1, complex128, M = 1e5 uniform points, 256x256, eps = 1e-9
  c3  3-D type 1, M = 5e7 uniform points, 256^3: eps 1e-4 and 1e-6 with
      complex64 values, 1e-9 with complex128 (FP64 path); error against the
      FP64 direct NUDFT on 10,000 sampled output modes
  c4  differentiable 2-D type 2, 512x512, M = 2e6, eps = 1e-6, complex64 and
      complex128: forward + backward (grid and coordinate gradients)

Inputs are synthetic (numpy/torch generators with fixed seeds), resident on
the GPU before timing.  Prints one JSON object per configuration.
python -m ducc_amd.configs_bench [--which c1,c3,c4] [--m3 50000000]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def _timed(fn, reps=3):
    import torch
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def _l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def c1():
    import torch
    from ducc_amd import nufft
    rng = np.random.default_rng(101)
    M, n, eps = 100_000, 256, 1e-9
    c = rng.uniform(-np.pi, np.pi, (M, 2))
    p = (rng.standard_normal(M) + 1j * rng.standard_normal(M)) * np.sqrt(0.5)
    ct, pt = torch.from_numpy(c).cuda(), torch.from_numpy(p).cuda()
    out = torch.empty((n, n), dtype=torch.complex128, device="cuda")
    info = {}
    t = _timed(lambda: nufft.nu2u(points=pt, coord=ct, forward=True,
                                  epsilon=eps, out=out, info=info))
    sel = rng.choice(n * n, 2000, replace=False)
    ref = nufft.nudft_direct(1, c, p, (n, n), True, sel)
    err = _l2(out.cpu().numpy().ravel()[sel], ref)
    return dict(config="c1", M=M, grid=[n, n], eps=eps, dtype="complex128",
                seconds=t, points_per_s=M / t, err_vs_direct=err,
                params=info)


def c3(M):
    import torch
    from ducc_amd import nufft
    g = torch.Generator(device="cuda").manual_seed(103)
    n = 256
    c = ((torch.rand((M, 3), generator=g, device="cuda", dtype=torch.float64)
          - 0.5) * (2 * math.pi)).contiguous()
    p128 = torch.randn(M, generator=g, device="cuda", dtype=torch.complex128)
    rng = np.random.default_rng(103)
    sel = rng.choice(n ** 3, 10000, replace=False)
    c_np = c.cpu().numpy()
    p_np = p128.cpu().numpy()
    ref = nufft.nudft_direct(1, c_np, p_np, (n, n, n), True, sel)
    res = []
    for eps, dt in ((1e-4, torch.complex64), (1e-6, torch.complex64),
                    (1e-9, torch.complex128)):
        pt = p128.to(dt)
        out = torch.empty((n, n, n), dtype=dt, device="cuda")
        info = {}
        t = _timed(lambda: nufft.nu2u(points=pt, coord=c, forward=True,
                                      epsilon=eps, out=out, info=info),
                   reps=2)
        got = out.cpu().numpy().ravel()[sel].astype(np.complex128)
        res.append(dict(config="c3", M=M, grid=[n, n, n], eps=eps,
                        dtype=str(dt).replace("torch.", ""), seconds=t,
                        points_per_s=M / t, err_vs_direct=_l2(got, ref),
                        params=info))
        del out
    return res


def c4():
    import torch
    from ducc_amd import torch_nufft as tn
    res = []
    rng = np.random.default_rng(104)
    M, n, eps = 2_000_000, 512, 1e-6
    c0 = rng.uniform(-np.pi, np.pi, (M, 2))
    g0 = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) \
        * np.sqrt(0.5)
    t0 = (rng.standard_normal(M) + 1j * rng.standard_normal(M)) * np.sqrt(0.5)
    for cdt, vdt in ((torch.float32, torch.complex64),
                     (torch.float64, torch.complex128)):
        c = torch.tensor(c0, device="cuda", dtype=torch.float64)
        gg = torch.tensor(g0, device="cuda", dtype=vdt, requires_grad=True)
        cc = c.clone().requires_grad_(True)
        tg = torch.tensor(t0, device="cuda", dtype=vdt)

        def step():
            gg.grad = None
            cc.grad = None
            L = (tn.u2nu(gg, cc, True, eps=eps) - tg).abs().pow(2).sum()
            L.backward()
        t = _timed(step)
        res.append(dict(config="c4", M=M, grid=[n, n], eps=eps,
                        dtype=str(vdt).replace("torch.", ""),
                        seconds_fwd_bwd=t, points_per_s=M / t))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--which", default="c1,c3,c4")
    ap.add_argument("--m3", type=int, default=50_000_000)
    a = ap.parse_args()
    out = []
    for w in a.which.split(","):
        if w == "c1":
            r = [c1()]
        elif w == "c3":
            r = c3(a.m3)
        elif w == "c4":
            r = c4()
        else:
            continue
        for x in r:
            print(json.dumps(x), flush=True)
            out.append(x)
    return out


if __name__ == "__main__":
    main()
