"""2-D float64 coordinates are read with one 16-byte load when the array is
16-byte aligned; a device view that starts 8 bytes into its buffer takes
the per-axis loads.  Both give the same transforms."""
import numpy as np
import pytest

import oracle.core as O

pytestmark = pytest.mark.gpu


def test_misaligned_device_coordinates(gpu):
    import torch
    from ducc_amd import nufft
    rng = np.random.default_rng(31)
    M, shape, eps = 20000, (96, 80), 1e-9
    c = rng.uniform(-np.pi, np.pi, (M, 2))
    p = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    g = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    buf = torch.zeros(2 * M + 1, dtype=torch.float64, device="cuda")
    buf[1:] = torch.from_numpy(c.ravel()).cuda()
    c_mis = buf[1:].view(M, 2)
    assert c_mis.data_ptr() % 16 == 8
    c_al = torch.from_numpy(c).cuda()
    assert c_al.data_ptr() % 16 == 0
    pt = torch.from_numpy(p).cuda()
    gt = torch.from_numpy(g).cuda()
    res = {}
    for name, cc in (("aligned", c_al), ("misaligned", c_mis)):
        o1 = torch.empty(shape, dtype=torch.complex128, device="cuda")
        o2 = torch.empty(M, dtype=torch.complex128, device="cuda")
        nufft.nu2u(points=pt, coord=cc, forward=True, epsilon=eps, out=o1)
        nufft.u2nu(grid=gt, coord=cc, forward=False, epsilon=eps, out=o2)
        res[name] = (o1.cpu().numpy(), o2.cpu().numpy())
    a, b = res["aligned"], res["misaligned"]
    # same parameters and sort; only the atomic summation order may differ
    assert O.l2error(a[0], b[0]) < 1e-13
    assert O.l2error(a[1], b[1]) < 1e-13
    sel = rng.choice(96 * 80, 64, replace=False)
    ref = O.nudft_type1(c, p, shape, True, modes=sel)
    assert O.l2error(b[0].ravel()[sel], ref) < 2 * eps
