"""The GPU direct FP64 NUDFT (k_nudft through dmx_nudft) against the
oracle's brute-force sums.  The GPU sum feeds the chooser's accuracy spot
check, the bench's error columns and the full-size configuration tests, so
it is pinned here on its own: types 1 and 2, 1-3 D, both mode orders,
non-2pi periodicity, and both launch shapes (one chunk of the inner sum per
output, and the split sum whose chunks are combined with atomics)."""
import numpy as np
import pytest

from oracle import core as O

pytestmark = pytest.mark.gpu


def _nudft():
    from ducc_amd import nufft
    return nufft.nudft_direct


@pytest.mark.parametrize("ndim,shape", [(1, (37,)), (2, (12, 9)),
                                        (3, (6, 5, 4))])
@pytest.mark.parametrize("M", [300, 5000])  # 300: one chunk; 5000: split
@pytest.mark.parametrize("fft_order", [False, True])
@pytest.mark.parametrize("forward", [True, False])
def test_type1_vs_oracle(gpu, ndim, shape, M, fft_order, forward):
    rng = np.random.default_rng(1000 + M + ndim)
    per = np.array([2 * np.pi, 3.7, 11.0][:ndim])
    c = rng.uniform(-1, 1, (M, ndim)) * per
    p = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    N = int(np.prod(shape))
    sel = rng.choice(N, size=min(N, 50), replace=False)
    got = _nudft()(1, c, p, shape, forward, sel, periodicity=per,
                   fft_order=fft_order)
    ref = O.nudft_type1(c, p, shape, forward, period=per,
                        fft_order=fft_order, modes=sel)
    assert O.l2error(got, ref) < 1e-12


@pytest.mark.parametrize("ndim,shape", [(1, (37,)), (2, (12, 9)),
                                        (2, (40, 33)), (3, (6, 5, 4)),
                                        (3, (10, 9, 8))])
@pytest.mark.parametrize("fft_order", [False, True])
@pytest.mark.parametrize("forward", [True, False])
def test_type2_vs_oracle(gpu, ndim, shape, fft_order, forward):
    # the inner sum runs over the modes: the small shapes take one chunk,
    # the larger ones (>= 512 modes) the split sum
    rng = np.random.default_rng(2000 + int(np.prod(shape)))
    per = np.array([2 * np.pi, 5.0, 0.9][:ndim])
    M = 400
    c = rng.uniform(-1, 1, (M, ndim)) * per
    g = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    sel = rng.choice(M, size=60, replace=False)
    got = _nudft()(2, c, g, shape, forward, sel, periodicity=per,
                   fft_order=fft_order)
    ref = O.nudft_type2(c, g, forward, period=per, fft_order=fft_order,
                        pidx=sel)
    assert O.l2error(got, ref) < 1e-12


def test_type1_matches_closed_form_single_point(gpu):
    """M = 1: out_k = c exp(isign i k x), independent of the oracle."""
    x = np.array([[0.7, -2.1]])
    v = np.array([0.3 - 1.2j])
    shape = (8, 6)
    sel = np.arange(48)
    got = _nudft()(1, x, v, shape, True, sel)
    k0 = np.arange(8) - 4
    k1 = np.arange(6) - 3
    ref = v[0] * np.exp(-1j * (k0[:, None] * x[0, 0] + k1[None, :] * x[0, 1]))
    assert np.max(np.abs(got - ref.ravel())) < 1e-14
