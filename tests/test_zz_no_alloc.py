"""dmx_nufft_noalloc, the allocation-free variant of dmx_nu2u / dmx_u2nu,
called through the C ABI and compared with the allocating entry points on
the same device inputs: two calls in a row per case, the second with a
smaller input (the static scratch keeps the first call's contents, so the
second must clear what it uses).  Both paths run the same kernels and take
the same parameters; only the FFT twiddles are computed on the device
instead of the host (last-bit differences), so the results agree to a
small multiple of the rounding error of the compute precision."""
import ctypes as ct

import numpy as np
import pytest
import torch

from oracle import core as O

pytestmark = pytest.mark.gpu


def _arr(t):
    from ducc_amd import _lib
    return _lib.describe(t)[0]


def _noalloc(type_, uni, coord, pts, forward, eps, per=None, fft_order=0,
             opt=None):
    from ducc_amd import _lib
    nd = coord.shape[1]
    per = np.full(nd, 2 * np.pi) if per is None else np.asarray(per, float)
    u, c, p = _arr(uni), _arr(coord), _arr(pts)
    inf = _lib.DmxPlanInfo()
    rc = _lib.lib().dmx_nufft_noalloc(
        type_, ct.byref(u), ct.byref(c), ct.byref(p), int(forward),
        float(eps), 1.2, 2.51, per.ctypes.data, int(fft_order),
        ct.byref(opt) if opt is not None else None, ct.byref(inf))
    return rc, _lib.lib().dmx_last_error().decode(), inf.as_dict()


def _case(seed, nd, M, shape, cdt):
    rng = np.random.default_rng(seed)
    c = torch.from_numpy(rng.uniform(-np.pi, np.pi, (M, nd))).cuda()
    p = torch.from_numpy(rng.standard_normal(M) +
                         1j * rng.standard_normal(M)).cuda().to(cdt)
    g = torch.from_numpy(rng.standard_normal(shape) +
                         1j * rng.standard_normal(shape)).cuda().to(cdt)
    return c, p, g


# (ndim, (large M, large shape), (small M, small shape), eps, dtype)
CASES = [
    (2, (300_000, (256, 200)), (50_000, (100, 90)), 1e-5, torch.complex64),
    (2, (200_000, (160, 160)), (20_000, (64, 48)), 1e-11, torch.complex128),
    (1, (100_000, (5000,)), (3_000, (777,)), 1e-6, torch.complex128),
    (3, (150_000, (48, 40, 32)), (10_000, (16, 20, 12)), 1e-4,
     torch.complex64),
]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("type_", [1, 2])
def test_noalloc_matches_allocating_entry(gpu, case, type_):
    from ducc_amd import nufft
    nd, big, small, eps, cdt = CASES[case]
    tol = 1e-12 if cdt == torch.complex128 else 2e-6
    for k, (M, shape) in enumerate((big, small)):
        print(f"noalloc case {case} type {type_} call {k}: M={M} "
              f"shape={shape}", flush=True)
        c, p, g = _case(100 * case + k, nd, M, shape, cdt)
        info = {}
        if type_ == 1:
            ref = torch.empty(shape, dtype=cdt, device="cuda")
            nufft.nu2u(points=p, coord=c, forward=True, epsilon=eps, out=ref,
                       info=info)
            got = torch.full(shape, 7.0, dtype=cdt, device="cuda")
            rc, msg, inf = _noalloc(1, got, c, p, True, eps)
        else:
            ref = nufft.u2nu(grid=g, coord=c, forward=False, epsilon=eps,
                             info=info)
            got = torch.full((M,), 7.0, dtype=cdt, device="cuda")
            rc, msg, inf = _noalloc(2, g, c, got, False, eps)
        torch.cuda.synchronize()
        assert rc == 0, msg
        assert (inf["kernel_idx"], inf["nover"]) == (info["kernel_idx"],
                                                     info["nover"])
        e = O.l2error(got.cpu().numpy(), ref.cpu().numpy())
        assert e <= tol, (k, e)


def test_noalloc_on_a_stream_async(gpu):
    """DMX_ASYNC on a given stream: the call returns once queued; work
    queued after it on the same stream sees the result."""
    from ducc_amd import _lib, nufft
    c, p, g = _case(7, 2, 100_000, (128, 96), torch.complex128)
    ref = nufft.u2nu(grid=g, coord=c, forward=True, epsilon=1e-9)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        out = torch.zeros(len(p), dtype=torch.complex128, device="cuda")
        opt = _lib.options(stream=side.cuda_stream)
        rc, msg, _ = _noalloc(2, g, c, out, True, 1e-9, opt=opt)
        assert rc == 0, msg
        doubled = out * 2
    side.synchronize()
    assert O.l2error(doubled.cpu().numpy(), 2 * ref.cpu().numpy()) < 1e-12


def test_noalloc_rejects_too_large_and_host_arrays(gpu):
    """A request beyond the static scratch, host arrays and the learned
    chooser are refused with a message, before any work is queued."""
    from ducc_amd import _lib
    c, p, g = _case(8, 2, 4_000_000, (16, 16), torch.complex128)
    out = torch.empty((4096, 4096), dtype=torch.complex128, device="cuda")
    rc, msg, _ = _noalloc(1, out, c, p, True, 1e-9)
    assert rc != 0 and "static scratch" in msg, msg
    cs = c[:1000].cpu().numpy()
    ps = p[:1000].cpu().numpy()
    rc, msg, _ = _noalloc(1, torch.empty((16, 16), dtype=torch.complex128,
                                         device="cuda"), cs, ps, True, 1e-9)
    assert rc != 0 and "device arrays only" in msg
    opt = _lib.options()
    opt.chooser = 1  # learned
    rc, msg, _ = _noalloc(1, torch.empty((16, 16), dtype=torch.complex128,
                                         device="cuda"), c[:1000], p[:1000],
                          True, 1e-9, opt=opt)
    assert rc != 0 and "learned chooser" in msg
