"""Calls on torch GPU tensors return once their work is queued on torch's
current stream (DMX_ASYNC): results are ordered on that stream like torch
operations, the library frees a call's buffers only after its kernels have
finished, and host arrays still return completed results."""
import numpy as np
import pytest

import oracle.core as O

pytestmark = pytest.mark.gpu


def _inputs(seed, M=30000, shape=(120, 96)):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-np.pi, np.pi, (M, 2))
    p = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    g = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return c, p, g


def test_back_to_back_calls_match_host_results(gpu):
    import torch
    from ducc_amd import nufft
    outs, refs = [], []
    for k in range(6):
        c, p, g = _inputs(40 + k)
        ct, pt, gt = (torch.from_numpy(x).cuda() for x in (c, p, g))
        o1 = torch.empty(g.shape, dtype=torch.complex128, device="cuda")
        o2 = torch.empty(len(p), dtype=torch.complex128, device="cuda")
        nufft.nu2u(points=pt, coord=ct, forward=True, epsilon=1e-9, out=o1)
        nufft.u2nu(grid=gt, coord=ct, forward=False, epsilon=1e-9, out=o2)
        outs.append((o1, o2))
        r1 = nufft.nu2u(points=p, coord=c, forward=True, epsilon=1e-9,
                        out=np.empty(g.shape, np.complex128))
        r2 = nufft.u2nu(grid=g, coord=c, forward=False, epsilon=1e-9)
        refs.append((r1, r2))
    torch.cuda.synchronize()
    for (o1, o2), (r1, r2) in zip(outs, refs):
        assert O.l2error(o1.cpu().numpy(), r1) < 1e-13
        assert O.l2error(o2.cpu().numpy(), r2) < 1e-13


def test_stream_order_on_a_side_stream(gpu):
    import torch
    from ducc_amd import nufft
    c, p, g = _inputs(50)
    ref = nufft.u2nu(grid=g, coord=c, forward=True, epsilon=1e-9)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        ct, gt = torch.from_numpy(c).cuda(), torch.from_numpy(g).cuda()
        out = torch.empty(len(p), dtype=torch.complex128, device="cuda")
        nufft.u2nu(grid=gt, coord=ct, forward=True, epsilon=1e-9, out=out)
        # overwrite the inputs in stream order right after the call: the
        # transform has read them first
        gt.zero_()
        ct.add_(1.0)
        doubled = out * 2
    side.synchronize()
    assert O.l2error(doubled.cpu().numpy(), 2 * ref) < 1e-13


def test_plan_on_device_is_stream_ordered(gpu):
    import torch
    from ducc_amd import nufft
    c, p, g = _inputs(60)
    ct = torch.from_numpy(c).cuda()
    pl = nufft.plan(nu2u=False, coord=ct, grid_shape=g.shape, epsilon=1e-9)
    outs = []
    for k in range(4):
        gk = torch.from_numpy(g * (k + 1)).cuda()
        o = torch.empty(len(p), dtype=torch.complex128, device="cuda")
        pl.u2nu(grid=gk, forward=True, out=o)
        outs.append(o)
    torch.cuda.synchronize()
    ref = nufft.u2nu(grid=g, coord=c, forward=True, epsilon=1e-9)
    for k, o in enumerate(outs):
        assert O.l2error(o.cpu().numpy(), (k + 1) * ref) < 1e-8


@pytest.mark.parametrize("kind", ["u2nu", "nu2u"])
def test_inputs_written_just_before_the_call(gpu, kind):
    """The grid side of an unplanned call runs on a second stream beside the
    index build: it must still see inputs written on the caller's stream
    right before the call (large inputs, so that a missed wait would read
    the old contents)."""
    import torch
    from ducc_amd import nufft
    # large enough for the second stream (M >= 2^20, oversampled grid of
    # at least 2^20 cells)
    c, p, g = _inputs(51, M=1_100_000, shape=(600, 500))
    if kind == "u2nu":
        ref = nufft.u2nu(grid=g, coord=c, forward=True, epsilon=1e-9)
    else:
        ref = nufft.nu2u(points=p, coord=c, forward=True, epsilon=1e-9,
                         out=np.empty(g.shape, np.complex128))
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        ct = torch.from_numpy(c).cuda()
        gsrc = torch.from_numpy(g).cuda()
        psrc = torch.from_numpy(p).cuda()
        gt = torch.zeros_like(gsrc)
        pt = torch.zeros_like(psrc)
        for _ in range(3):  # several kernels queued ahead of the call
            gt.copy_(gsrc)
            pt.copy_(psrc)
        if kind == "u2nu":
            out = torch.empty(len(p), dtype=torch.complex128, device="cuda")
            nufft.u2nu(grid=gt, coord=ct, forward=True, epsilon=1e-9,
                       out=out)
        else:
            out = torch.empty(g.shape, dtype=torch.complex128, device="cuda")
            nufft.nu2u(points=pt, coord=ct, forward=True, epsilon=1e-9,
                       out=out)
        res = out.clone()
    side.synchronize()
    assert O.l2error(res.cpu().numpy(), ref) < 1e-13
