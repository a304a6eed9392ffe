"""Autograd of the GPU NUFFT: the u2nu backward is the hand-written adjoint
(type 1 with the opposite sign) plus the kernel-derivative interpolation for
coordinate gradients.  Checked against a dense torch reference (float64 CPU
autograd of the explicit sum), torch.autograd.gradcheck, FP64 central
differences (h = 1e-6) at the 512x512, M = 2e6, eps = 1e-6 configuration,
and the exact derivative of the direct sum."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def dense_u2nu(grid, coord, forward, period=2 * np.pi):
    """explicit sum with torch ops (CPU, float64) - differentiable"""
    shape = grid.shape
    ks = [torch.arange(n, dtype=torch.float64) - n // 2 for n in shape]
    mesh = torch.meshgrid(*ks, indexing="ij")
    kk = torch.stack([m.reshape(-1) for m in mesh], 1)  # (N, d)
    s = -1.0 if forward else 1.0
    ph = s * (2 * np.pi / period) * (coord @ kk.T)  # (M, N)
    A = torch.polar(torch.ones_like(ph), ph)
    return A @ grid.reshape(-1)


def dense_nu2u(points, coord, shape, forward, period=2 * np.pi):
    ks = [torch.arange(n, dtype=torch.float64) - n // 2 for n in shape]
    mesh = torch.meshgrid(*ks, indexing="ij")
    kk = torch.stack([m.reshape(-1) for m in mesh], 1)
    s = -1.0 if forward else 1.0
    ph = s * (2 * np.pi / period) * (kk @ coord.T)  # (N, M)
    A = torch.polar(torch.ones_like(ph), ph)
    return (A @ points).reshape(shape)


@pytest.mark.parametrize("shape", [(24,), (12, 10), (6, 5, 7)])
@pytest.mark.parametrize("forward", [True, False])
def test_u2nu_grads_vs_dense(gpu, shape, forward):
    from ducc_amd import torch_nufft as tn
    rng = np.random.default_rng(1)
    M = 40
    nd = len(shape)
    c0 = rng.uniform(-np.pi, np.pi, (M, nd))
    g0 = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    t0 = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    # dense CPU reference
    c = torch.tensor(c0, requires_grad=True)
    g = torch.tensor(g0, requires_grad=True)
    t = torch.tensor(t0)
    L = (dense_u2nu(g, c, forward) - t).abs().pow(2).sum()
    L.backward()
    # GPU
    cg = torch.tensor(c0, device="cuda", requires_grad=True)
    gg = torch.tensor(g0, device="cuda", requires_grad=True)
    tg = torch.tensor(t0, device="cuda")
    Lg = (tn.u2nu(gg, cg, forward, eps=1e-12) - tg).abs().pow(2).sum()
    Lg.backward()
    assert abs(Lg.item() - L.item()) <= 1e-9 * L.item()
    gref = g.grad.numpy()
    assert np.linalg.norm(gg.grad.cpu().numpy() - gref) <= 1e-9 * np.linalg.norm(gref)
    cref = c.grad.numpy()
    assert np.linalg.norm(cg.grad.cpu().numpy() - cref) <= 1e-7 * np.linalg.norm(cref)


@pytest.mark.parametrize("shape", [(20,), (9, 11), (5, 6, 4)])
def test_nu2u_grads_vs_dense(gpu, shape):
    from ducc_amd import torch_nufft as tn
    rng = np.random.default_rng(2)
    M = 30
    nd = len(shape)
    c0 = rng.uniform(-np.pi, np.pi, (M, nd))
    p0 = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    t0 = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    c = torch.tensor(c0, requires_grad=True)
    p = torch.tensor(p0, requires_grad=True)
    L = (dense_nu2u(p, c, shape, True) - torch.tensor(t0)).abs().pow(2).sum()
    L.backward()
    cg = torch.tensor(c0, device="cuda", requires_grad=True)
    pg = torch.tensor(p0, device="cuda", requires_grad=True)
    Lg = (tn.nu2u(pg, cg, shape, True, eps=1e-12) -
          torch.tensor(t0, device="cuda")).abs().pow(2).sum()
    Lg.backward()
    assert abs(Lg.item() - L.item()) <= 1e-9 * L.item()
    assert np.allclose(pg.grad.cpu().numpy(), p.grad.numpy(), rtol=0,
                       atol=1e-9 * np.abs(p.grad.numpy()).max())
    assert np.allclose(cg.grad.cpu().numpy(), c.grad.numpy(), rtol=0,
                       atol=1e-7 * np.abs(c.grad.numpy()).max())


def test_gradcheck_small(gpu):
    from ducc_amd import torch_nufft as tn
    rng = np.random.default_rng(3)
    c = torch.tensor(rng.uniform(-np.pi, np.pi, (8, 2)), device="cuda",
                     requires_grad=True)
    g = torch.tensor(rng.standard_normal((6, 5)) +
                     1j * rng.standard_normal((6, 5)), device="cuda",
                     requires_grad=True)
    f = lambda gg, cc: tn.u2nu(gg, cc, True, eps=1e-13)
    assert torch.autograd.gradcheck(f, (g, c), eps=1e-6, atol=1e-6,
                                    rtol=1e-5)


def test_c4_central_differences(gpu):
    """512x512 grid, M = 2e6 uniform points, eps = 1e-6, complex128:
    loss = sum |u2nu(g) - t|^2; 64 gradient entries w.r.t. the grid
    (re and im of 32 modes) and 64 w.r.t. coordinates (32 points x 2 axes)
    against FP64 central differences with h = 1e-6."""
    from ducc_amd import nufft
    from ducc_amd import torch_nufft as tn
    rng = np.random.default_rng(104)
    M, n, eps, h = 2_000_000, 512, 1e-6, 1e-6
    c0 = rng.uniform(-np.pi, np.pi, (M, 2))
    g0 = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) \
        * np.sqrt(0.5)
    t0 = (rng.standard_normal(M) + 1j * rng.standard_normal(M)) * np.sqrt(0.5)
    cg = torch.tensor(c0, device="cuda", requires_grad=True)
    gg = torch.tensor(g0, device="cuda", requires_grad=True)
    tg = torch.tensor(t0, device="cuda")
    out = tn.u2nu(gg, cg, True, eps=eps)
    L = (out - tg).abs().pow(2).sum()
    L.backward()
    ggrad = gg.grad.cpu().numpy()
    cgrad = cg.grad.cpu().numpy()

    def fwd(g, c):
        return nufft.u2nu(grid=g, coord=c, forward=True, epsilon=eps)

    def dloss(op, om):
        # elementwise difference of the loss terms: no cancellation of L
        return float(torch.sum((op - tg).abs() ** 2 -
                               (om - tg).abs() ** 2).item())

    g_t = torch.tensor(g0, device="cuda")
    c_t = torch.tensor(c0, device="cuda")
    errs = []
    for q in rng.choice(n * n, 32, replace=False):
        i, j = divmod(int(q), n)
        for part, unit in (("re", 1.0), ("im", 1j)):
            gp = g_t.clone()
            gm = g_t.clone()
            gp[i, j] += unit * h
            gm[i, j] -= unit * h
            fd = dloss(fwd(gp, c_t), fwd(gm, c_t)) / (2 * h)
            an = ggrad[i, j].real if part == "re" else ggrad[i, j].imag
            errs.append(abs(fd - an) / max(abs(an), 1e-3 * np.abs(ggrad).max()))
    assert max(errs) < 1e-4, max(errs)
    errs = []
    for p in rng.choice(M, 32, replace=False):
        for ax in range(2):
            cp = c_t.clone()
            cm = c_t.clone()
            cp[p, ax] += h
            cm[p, ax] -= h
            fd = dloss(fwd(g_t, cp), fwd(g_t, cm)) / (2 * h)
            an = cgrad[p, ax]
            errs.append(abs(fd - an) / max(abs(an), 1e-3 * np.abs(cgrad).max()))
    # a point within ~2e-4 cells of a kernel piece boundary sees the small
    # jump of the piecewise polynomial: allow one such entry
    errs = sorted(errs)
    assert errs[-2] < 1e-4, errs[-5:]
    # coordinate gradient against the exact derivative of the direct sum
    sel = rng.choice(M, 16, replace=False)
    k = np.arange(n) - n // 2
    r = (out.detach().cpu().numpy() - t0)
    for p in sel:
        e0 = np.exp(-1j * k * c0[p, 0])
        e1 = np.exp(-1j * k * c0[p, 1])
        D0 = np.sum(g0 * (-1j * k)[:, None] * e0[:, None] * e1[None, :])
        D1 = np.sum(g0 * e0[:, None] * ((-1j * k) * e1)[None, :])
        ex = [np.real(2 * r[p] * np.conj(D0)), np.real(2 * r[p] * np.conj(D1))]
        # the NUFFT error of a sum sum_k a_k e^{ikx} is ~eps*||a||_2; the
        # derivative's terms are g*k
        scale = [np.linalg.norm(g0 * k[:, None]), np.linalg.norm(g0 * k[None, :])]
        for ax in range(2):
            assert abs(cgrad[p, ax] - ex[ax]) <= 100 * eps * 2 * abs(r[p]) * \
                scale[ax], (p, ax, cgrad[p, ax], ex[ax])


def test_complex64_autograd_close_to_complex128(gpu):
    from ducc_amd import torch_nufft as tn
    rng = np.random.default_rng(5)
    M, shape = 20000, (64, 48)
    c0 = rng.uniform(-np.pi, np.pi, (M, 2))
    g0 = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    res = []
    for dt, cdt in ((torch.complex128, torch.float64),
                    (torch.complex64, torch.float32)):
        cg = torch.tensor(c0, device="cuda", dtype=cdt, requires_grad=True)
        gg = torch.tensor(g0, device="cuda", dtype=dt, requires_grad=True)
        L = tn.u2nu(gg, cg, True, eps=1e-5).abs().pow(2).sum()
        L.backward()
        res.append((gg.grad.cpu().numpy().astype(np.complex128),
                    cg.grad.cpu().numpy().astype(np.float64)))
    (g64, c64), (g32, c32) = res
    assert np.linalg.norm(g32 - g64) <= 1e-4 * np.linalg.norm(g64)
    assert np.linalg.norm(c32 - c64) <= 1e-3 * np.linalg.norm(c64)
