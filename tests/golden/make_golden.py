"""Write the golden fixtures of this directory from the reference ducc0,
compiled from its own sources into oracle/_ref (make -C oracle ref).

Fixtures (inputs and ducc0's outputs, all seeded, small):
  params.json     ducc0 findNufftParameters (row, oversampled shape) for
                  random requests, with the SIMD width of the build
  kernels.npz     piecewise-polynomial coefficients and correction factors
                  of selected kernel-table rows
  sort.npz        ducc0's processing order (coord_idx) for a few geometries
  nufft.npz       nu2u / u2nu results for 1-3 D, both precisions,
                  forward/backward, fft_order, both periodicities
Run: python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import ref as R  # noqa: E402
from oracle import core as O  # noqa: E402


def params():
    rng = np.random.default_rng(2024)
    reqs = []
    for _ in range(400):
        nd = int(rng.integers(1, 4))
        hi = {1: 100000, 2: 3000, 3: 300}[nd]
        shape = [int(x) for x in np.exp(rng.uniform(np.log(1), np.log(hi),
                                                     nd))]
        single = bool(rng.integers(0, 2))
        eps = float(10 ** rng.uniform(-14 if not single else -6.5, -1))
        M = int(10 ** rng.uniform(0, 8))
        grid = bool(rng.integers(0, 2))
        nth = int(rng.integers(1, 17))
        smin, smax = [(1.2, 2.51), (1.1, 2.6), (1.25, 2.0)][
            int(rng.integers(0, 3))]
        try:
            k, nov = R.find_params(eps, shape, M, grid, single, nth, smin,
                                   smax)
        except RuntimeError:
            k, nov = -1, []
        reqs.append(dict(ndim=nd, shape=shape, single=single, eps=eps, M=M,
                         gridding=grid, nthreads=nth, sigma_min=smin,
                         sigma_max=smax, kidx=k, nover=list(nov)))
    best = {f"{nd}{'s' if s else 'd'}": R.best_epsilon(nd, s)
            for nd in (1, 2, 3) for s in (True, False)}
    out = dict(vlen_single=R.vlen(True), vlen_double=R.vlen(False),
               requests=reqs, best_epsilon=best)
    with open(os.path.join(HERE, "params.json"), "w") as f:
        json.dump(out, f, indent=0)


def kernels():
    db = O.load_db()
    rows = [0, 37, 135, 200, 486, 700, 835, 900, 1186, 1200, 1424]
    d = {}
    for k in rows:
        W = int(db["W"][k])
        d[f"coeff_{k}"] = R.kernel_coeffs(k, W)
        d[f"corfac_{k}"] = R.corfac(k, 300, 1 / 517.)
    d["rows"] = np.array(rows)
    np.savez_compressed(os.path.join(HERE, "kernels.npz"), **d)


def sort():
    rng = np.random.default_rng(77)
    d = {}
    cases = [(1, (1000,), 1e-5, 0), (2, (200, 300), 1e-5, 1),
             (2, (128, 128), 1e-9, 0), (3, (40, 50, 30), 1e-4, 1),
             (3, (32, 32, 32), 1e-8, 0)]
    for i, (nd, shape, eps, single) in enumerate(cases):
        M = 5000
        c = rng.uniform(-np.pi, np.pi, (M, nd))
        idx, nover, supp = R.sort_index(c, shape, eps, True, single)
        d[f"coord_{i}"] = c
        d[f"idx_{i}"] = idx
        d[f"meta_{i}"] = np.array([nd, single, supp] + list(nover),
                                  dtype=np.int64)
        d[f"shape_{i}"] = np.array(shape)
        d[f"eps_{i}"] = np.array(eps)
    np.savez_compressed(os.path.join(HERE, "sort.npz"), **d)


def nufft():
    rng = np.random.default_rng(4242)
    d = {}
    cases = []
    shapes = [(37,), (256,), (20, 33), (64, 48), (1, 21), (12, 10, 14),
              (16, 16, 16)]
    i = 0
    for shape in shapes:
        nd = len(shape)
        for single, eps in ((False, 1e-3), (False, 1e-6), (False, 1e-11),
                            (True, 1e-3), (True, 5e-6)):
            fwd = bool(rng.integers(0, 2))
            fo = bool(rng.integers(0, 2))
            per = 2 * np.pi if rng.integers(0, 2) else 1.0
            M = int(rng.integers(50, 400))
            c = (rng.random((M, nd)) - 0.5) * per
            p = rng.standard_normal(M) + 1j * rng.standard_normal(M)
            g = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
            cdt = np.float64
            vdt = np.complex64 if single else np.complex128
            o1 = R.nu2u(p.astype(vdt), c.astype(cdt), fwd, eps, shape,
                        nthreads=1, periodicity=per, fft_order=fo)
            o2 = R.u2nu(g.astype(vdt), c.astype(cdt), not fwd, eps,
                        nthreads=1, periodicity=per, fft_order=fo)
            d[f"coord_{i}"] = c
            d[f"points_{i}"] = p.astype(vdt)
            d[f"grid_{i}"] = g.astype(vdt)
            d[f"nu2u_{i}"] = o1
            d[f"u2nu_{i}"] = o2
            cases.append(dict(shape=list(shape), single=single, eps=eps,
                              forward=fwd, fft_order=fo, periodicity=per))
            i += 1
    d["cases"] = np.array(json.dumps(cases))
    np.savez_compressed(os.path.join(HERE, "nufft.npz"), **d)


if __name__ == "__main__":
    if not R.available():
        sys.exit("oracle/_ref/libducc_ref.so missing: run make -C oracle ref")
    params()
    kernels()
    sort()
    nufft()
    print("golden fixtures written to", HERE)
