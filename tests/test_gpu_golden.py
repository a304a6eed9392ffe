"""The GPU path against ducc0's own outputs (golden fixtures written from the
compiled reference by tests/golden/make_golden.py): the plain path must match
ducc0's nu2u / u2nu within epsilon (relative l2, misc_pymod.cc:129-170)."""
import json
import os

import numpy as np
import pytest

from oracle import core as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _cases():
    z = np.load(os.path.join(GOLD, "nufft.npz"))
    return z, json.loads(str(z["cases"]))


@pytest.mark.parametrize("i", range(35))
def test_matches_ducc0(gpu, i):
    from ducc_amd import nufft
    z, cases = _cases()
    cs = cases[i]
    c, p, g = z[f"coord_{i}"], z[f"points_{i}"], z[f"grid_{i}"]
    shape = tuple(cs["shape"])
    eps, fwd = cs["eps"], cs["forward"]
    kw = dict(periodicity=cs["periodicity"], fft_order=cs["fft_order"])
    out = np.empty(shape, p.dtype)
    nufft.nu2u(points=p, coord=c, forward=fwd, epsilon=eps, out=out, **kw)
    m = nufft.u2nu(grid=g, coord=c, forward=not fwd, epsilon=eps, **kw)
    # both sides are eps-accurate approximations of the same sums
    assert O.l2error(out, z[f"nu2u_{i}"]) <= 2 * eps
    assert O.l2error(m, z[f"u2nu_{i}"]) <= 2 * eps
    # and the exact sums confirm which side is off, if any
    ref = O.nudft_type1(c, p, shape, fwd, period=cs["periodicity"],
                        fft_order=cs["fft_order"])
    mul = {1: 4, 2: 2, 3: 1}[len(shape)]
    assert O.l2error(out, ref) <= mul * eps
