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


def test_numpy_plan_reads_torch_input_in_stream_order(gpu):
    """A plan made from host (numpy) coordinates, fed a device tensor that
    torch is still writing on its current stream: the planned call runs on
    that stream, so it reads the finished values."""
    import torch
    from ducc_amd import nufft
    c, p, g = _inputs(61, M=400_000, shape=(300, 256))
    pl = nufft.plan(nu2u=True, coord=c, grid_shape=g.shape, epsilon=1e-9)
    ref = nufft.nu2u(points=p, coord=c, forward=True, epsilon=1e-9,
                     out=np.empty(g.shape, np.complex128))
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        psrc = torch.from_numpy(p).cuda()
        pt = torch.zeros_like(psrc)
        big = torch.randn(3000, 3000, device="cuda", dtype=torch.float64)
        for _ in range(4):  # slow work queued ahead of the final write
            big = big @ big.T / 3000.0
        pt.copy_(psrc + 0 * big[0, 0])
        out = pl.nu2u(points=pt, forward=True)
        res = out.clone()
    side.synchronize()
    assert O.l2error(res.cpu().numpy(), ref) < 1e-13


def test_device_plan_called_under_a_side_stream(gpu):
    """A plan made from device coordinates on the default stream and called
    under another stream: the results are ordered on the calling stream
    (a consumer queued right after the call sees them), and a later call
    on the first stream waits for the plan's work on the side stream."""
    import torch
    from ducc_amd import nufft
    c, p, g = _inputs(62, M=400_000, shape=(300, 256))
    ref = nufft.u2nu(grid=g, coord=c, forward=True, epsilon=1e-9)
    ct = torch.from_numpy(c).cuda()
    pl = nufft.plan(nu2u=False, coord=ct, grid_shape=g.shape, epsilon=1e-9)
    gt = torch.from_numpy(g).cuda()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        out = pl.u2nu(grid=gt, forward=True)
        doubled = out * 2
    # back on the default stream: the same plan again, then a consumer
    out2 = pl.u2nu(grid=gt * 3, forward=True)
    tripled = out2.clone()
    torch.cuda.synchronize()
    assert O.l2error(doubled.cpu().numpy(), 2 * ref) < 1e-13
    assert O.l2error(tripled.cpu().numpy(), 3 * ref) < 1e-13


def test_plan_from_host_coords_then_other_stream_allocations(gpu):
    """Host coordinates are staged for a plan's index build, which runs
    asynchronously: the staging block must not be handed to other work
    before the build has read it.  A plan from numpy coordinates is
    followed at once by unplanned torch calls on a side stream (which
    allocate from the same pool and overwrite what they get); the plan's
    results must still match the direct sum."""
    import torch
    from ducc_amd import nufft
    rng = np.random.default_rng(63)
    M, shape = 600_000, (200, 180)
    c = rng.uniform(-np.pi, np.pi, (M, 2))
    p = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    pl = nufft.plan(nu2u=True, coord=c, grid_shape=shape, epsilon=1e-9)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        for k in range(3):
            cc = torch.from_numpy(rng.uniform(-np.pi, np.pi, (M, 2))).cuda()
            pp = torch.from_numpy(p).cuda()
            o = torch.empty(shape, dtype=torch.complex128, device="cuda")
            nufft.nu2u(points=pp, coord=cc, forward=True, epsilon=1e-9,
                       out=o)
    res = pl.nu2u(points=p, forward=True)
    side.synchronize()
    sel = rng.choice(shape[0] * shape[1], 200, replace=False)
    ref = O.nudft_type1(c, p, shape, True, modes=sel)
    assert O.l2error(res.ravel()[sel], ref) < 2e-9


def test_threads_with_their_own_plans(gpu):
    """Reentrancy: four host threads transform concurrently (the library
    releases no lock around GPU work; ctypes drops the GIL), each with its
    own plan and unplanned calls; every result matches its own reference."""
    import threading

    from ducc_amd import nufft
    errs = {}

    def work(k):
        c, p, g = _inputs(70 + k, M=150_000, shape=(128, 112))
        pl = nufft.plan(nu2u=False, coord=c, grid_shape=g.shape,
                        epsilon=1e-9)
        a = pl.u2nu(grid=g, forward=True)
        b = nufft.u2nu(grid=g, coord=c, forward=True, epsilon=1e-9)
        sel = np.arange(0, len(p), 997)
        ref = O.nudft_type2(c, g, True, pidx=sel)
        errs[k] = (O.l2error(a[sel], ref), O.l2error(b[sel], ref))

    ts = [threading.Thread(target=work, args=(k,)) for k in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert len(errs) == 4
    for k, (ea, eb) in errs.items():
        assert ea < 2e-9 and eb < 2e-9, (k, ea, eb)


def test_verbosity_prints_request_and_phase_table(gpu, capfd):
    """verbosity > 0 prints ducc0's request report (nufft.h:371-380) and the
    phase timing table (timers.h:117-160) for unplanned and planned calls;
    verbosity 0 prints nothing."""
    from ducc_amd import nufft
    c, p, g = _inputs(80)
    capfd.readouterr()
    nufft.nu2u(points=p, coord=c, forward=True, epsilon=1e-9,
               out=np.empty(g.shape, np.complex128))
    assert capfd.readouterr().out == ""
    nufft.nu2u(points=p, coord=c, forward=True, epsilon=1e-9,
               out=np.empty(g.shape, np.complex128), verbosity=1)
    out = capfd.readouterr().out
    assert "Nu2u:" in out and "oversampled grid=(" in out
    assert "Total wall clock time for nu2u" in out
    assert "+- spreading" in out and "+- building index" in out
    pl = nufft.plan(nu2u=False, coord=c, grid_shape=g.shape, epsilon=1e-9)
    capfd.readouterr()
    pl.u2nu(grid=g, forward=True, verbosity=2)
    out = capfd.readouterr().out
    assert "U2nu:" in out and "Total wall clock time for u2nu" in out
    assert "+- interpolation" in out
