"""GPU NUFFT parity against the CPU oracle (oracle/oracle.c, a restatement of
ducc0's algorithm in FP64) and the brute-force NUDFT, following the
reference's test strategy (python/test/test_nufft.py): l2 error vs the
explicit sum within 4/2/1 eps in 1/2/3-D, adjointness, plan == unplanned."""
import numpy as np
import pytest

from oracle import core as O

pytestmark = pytest.mark.gpu
pmp = pytest.mark.parametrize


def vdot(a, b):
    return np.vdot(np.asarray(a, np.complex128).ravel(),
                   np.asarray(b, np.complex128).ravel())


def data(rng, shape, M, period, single):
    ndim = len(shape)
    c = (rng.random((M, ndim)) - 0.5) * np.asarray(period)
    ms = rng.random(M) - 0.5 + 1j * (rng.random(M) - 0.5)
    g = rng.random(shape) - 0.5 + 1j * (rng.random(shape) - 0.5)
    if single:
        ms = ms.astype(np.complex64)
        g = g.astype(np.complex64)
    return c, ms, g


TOLMUL = {1: 4, 2: 2, 3: 1}

SHAPES = [(1,), (20,), (257,), (21, 32), (250, 257), (1, 21), (22, 35, 1),
          (22, 35, 21)]


@pmp("shape", SHAPES)
@pmp("eps,single", [(1e-1, False), (3e-5, False), (2e-13, False),
                    (1e-1, True), (3e-5, True)])
@pmp("forward", [True, False])
@pmp("periodicity", ["2pi", "unit"])
@pmp("fft_order", [False, True])
def test_vs_explicit(gpu, shape, eps, single, forward, periodicity,
                     fft_order):
    from ducc_amd import nufft
    ndim = len(shape)
    if ndim == 3 and eps < 1e-12:
        eps = 5e-13
    rng = np.random.default_rng(42)
    per = 2 * np.pi if periodicity == "2pi" else [1.0] * ndim
    M = 37
    c, ms, g = data(rng, shape, M, per, single)
    out = np.empty(shape, ms.dtype)
    nufft.nu2u(points=ms, coord=c, forward=forward, epsilon=eps, out=out,
               periodicity=per, fft_order=fft_order)
    ref = O.nudft_type1(c, ms, shape, forward, period=per,
                        fft_order=fft_order)
    assert O.l2error(out, ref) <= TOLMUL[ndim] * eps
    m2 = nufft.u2nu(grid=g, coord=c, forward=not forward, epsilon=eps,
                    periodicity=per, fft_order=fft_order)
    ref2 = O.nudft_type2(c, g, not forward, period=per, fft_order=fft_order)
    assert O.l2error(m2, ref2) <= TOLMUL[ndim] * eps
    # adjointness (test_nufft.py:68-72)
    d2 = out
    r = max(vdot(ms, ms).real, vdot(m2, m2).real, vdot(g, g).real,
            vdot(d2, d2).real)
    tol = 3e-5 * r if single else 2e-13 * r
    assert abs(vdot(ms, m2) - vdot(d2, g)) <= tol + 1e-300


@pmp("shape,M", [((1000,), 5000), ((128, 96), 20000), ((24, 30, 20), 20000)])
@pmp("eps", [1e-3, 1e-6, 1e-10])
def test_vs_oracle_restatement(gpu, shape, M, eps):
    """FP64 path against the CPU restatement with the same kernel and grid:
    both compute ducc0's approximation, so they agree to ~1e-13."""
    from ducc_amd import nufft
    rng = np.random.default_rng(7)
    c, ms, g = data(rng, shape, M, 2 * np.pi, False)
    info = {}
    out = np.empty(shape, np.complex128)
    nufft.nu2u(points=ms, coord=c, forward=True, epsilon=eps, out=out,
               precision="fp64", info=info)
    oinfo = {}
    ref = O.nu2u(c, ms, shape, True, eps, single=False, vlen=8, nthreads=8,
                 info=oinfo)
    assert info["kernel_idx"] == oinfo["kidx"]
    assert tuple(info["nover"]) == oinfo["nover"]
    assert O.l2error(out, ref) < 1e-12
    m2 = nufft.u2nu(grid=g, coord=c, forward=False, epsilon=eps,
                    precision="fp64")
    ref2 = O.u2nu(c, g, False, eps, single=False, vlen=8, nthreads=8)
    assert O.l2error(m2, ref2) < 1e-12


@pmp("shape,M", [((512,), 20000), ((300, 200), 50000), ((40, 36, 30), 30000)])
def test_fp32_path(gpu, shape, M):
    """FP32 compute (FP64 fold) meets eps against the FP64 NUDFT."""
    from ducc_amd import nufft
    rng = np.random.default_rng(11)
    c, ms, g = data(rng, shape, M, 2 * np.pi, True)
    eps = 1e-5
    info = {}
    out = np.empty(shape, np.complex64)
    nufft.nu2u(points=ms, coord=c, forward=True, epsilon=eps, out=out,
               info=info)
    assert not info["fp64"]
    sel = rng.choice(int(np.prod(shape)), 200, replace=False)
    ref = O.nudft_type1(c, ms, shape, True, modes=sel)
    assert O.l2error(out.ravel()[sel], ref) <= eps
    m2 = nufft.u2nu(grid=g, coord=c, forward=False, epsilon=eps)
    pidx = rng.choice(M, 200, replace=False)
    ref2 = O.nudft_type2(c, g, False, pidx=pidx)
    assert O.l2error(m2[pidx], ref2) <= eps


@pmp("shape", [(64,), (48, 40), (16, 18, 20)])
def test_plan_matches_unplanned(gpu, shape):
    from ducc_amd import nufft
    rng = np.random.default_rng(3)
    c, ms, g = data(rng, shape, 3000, 2 * np.pi, False)
    eps = 1e-9
    a = np.empty(shape, np.complex128)
    nufft.nu2u(points=ms, coord=c, forward=True, epsilon=eps, out=a,
               sigma_min=1.1, sigma_max=2.6)
    p = nufft.plan(nu2u=True, coord=c, grid_shape=shape, epsilon=eps)
    b = p.nu2u(points=ms, forward=True)
    assert O.l2error(a, b) < 1e-13
    b2 = p.nu2u(points=ms, forward=True)
    assert O.l2error(b, b2) < 1e-13
    q = nufft.plan(nu2u=False, coord=c, grid_shape=shape, epsilon=eps)
    m1 = q.u2nu(grid=g, forward=False)
    m2 = nufft.u2nu(grid=g, coord=c, forward=False, epsilon=eps,
                    sigma_min=1.1, sigma_max=2.6)
    assert O.l2error(m1, m2) < 1e-13


@pmp("shape,single,fft_order", [
    ((300,), False, False), ((9000,), False, True), ((9000,), True, False),
    ((48, 40), False, True), ((250, 257), False, False),
    ((250, 257), True, True), ((64, 5000), False, False),
    ((16, 18, 20), False, False), ((16, 18, 20), True, True)])
def test_u2nu_ignores_stale_grid(gpu, shape, single, fft_order):
    """Type 2 writes only the grid region its FFTs read (retained modes
    along axes whose FFT runs in LDS, zeros elsewhere): a grid buffer that
    still holds a type-1 result must not change the output."""
    from ducc_amd import nufft
    rng = np.random.default_rng(11)
    c, ms, g = data(rng, shape, 4000, 2 * np.pi, single)
    if single:
        c = c.astype(np.float32)
    eps = 1e-5 if single else 1e-10
    p = nufft.plan(nu2u=False, coord=c, grid_shape=shape, epsilon=eps,
                   fft_order=fft_order)
    m1 = p.u2nu(grid=g, forward=True)
    p.nu2u(points=ms, forward=False)
    m2 = p.u2nu(grid=g, forward=True)
    assert np.array_equal(m1, m2)
    gg = np.fft.fftshift(g) if fft_order else g
    pidx = rng.choice(4000, 300, replace=False)
    ref = O.nudft_type2(c.astype(np.float64), gg.astype(np.complex128), True,
                        pidx=pidx)
    assert O.l2error(m2[pidx], ref) <= TOLMUL[len(shape)] * eps


def test_coord_and_value_dtype_matrix(gpu):
    """coords {f64,f32} x values {c128,c64} (nufft_pymod.cc:70-94)"""
    from ducc_amd import nufft
    rng = np.random.default_rng(9)
    shape = (40, 30)
    c, ms, g = data(rng, shape, 500, 2 * np.pi, False)
    ref = O.nudft_type1(c.astype(np.float32).astype(np.float64), ms, shape,
                        True)
    for cd in (np.float64, np.float32):
        for vd in (np.complex128, np.complex64):
            out = np.empty(shape, vd)
            nufft.nu2u(points=ms.astype(vd), coord=c.astype(cd),
                       forward=True, epsilon=1e-5, out=out)
            assert out.dtype == vd
            tol = 2e-5 if cd == np.float64 else 1e-4
            assert O.l2error(out, ref) < tol


def test_strided_host_arrays(gpu):
    from ducc_amd import nufft
    rng = np.random.default_rng(12)
    shape = (30, 40)
    c, ms, g = data(rng, shape, 400, 2 * np.pi, False)
    cbig = np.zeros((400, 5))
    cbig[:, 1:5:2] = c
    cs = cbig[:, 1:5:2]
    gbig = np.zeros((60, 40), np.complex128)
    gbig[::2] = g
    gs = gbig[::2]
    m_ref = nufft.u2nu(grid=g, coord=c, forward=True, epsilon=1e-10)
    m_s = nufft.u2nu(grid=gs, coord=cs, forward=True, epsilon=1e-10)
    assert O.l2error(m_ref, m_s) < 1e-14
    obig = np.zeros((30, 80), np.complex128)
    nufft.nu2u(points=ms, coord=cs, forward=True, epsilon=1e-10,
               out=obig[:, ::2])
    o_ref = np.empty(shape, np.complex128)
    nufft.nu2u(points=ms, coord=c, forward=True, epsilon=1e-10, out=o_ref)
    # spreading accumulates with float atomics: summation order (and so the
    # last bits) varies from run to run
    assert O.l2error(obig[:, ::2], o_ref) < 1e-12
    assert np.all(obig[:, 1::2] == 0)


def test_errors(gpu):
    from ducc_amd import nufft
    c = np.zeros((10, 2))
    with pytest.raises(RuntimeError, match="not yet supported"):
        nufft.u2nu(grid=np.zeros((4, 4), np.float64), coord=c, forward=True,
                   epsilon=1e-5)
    with pytest.raises(RuntimeError, match="number of points mismatch"):
        nufft.nu2u(points=np.zeros(9, np.complex128), coord=c, forward=True,
                   epsilon=1e-5, out=np.zeros((4, 4), np.complex128))
    with pytest.raises(RuntimeError, match="no appropriate kernel found"):
        nufft.nu2u(points=np.zeros(10, np.complex128), coord=c, forward=True,
                   epsilon=1e-20, out=np.zeros((4, 4), np.complex128))
    with pytest.raises(RuntimeError, match="dimensionality mismatch"):
        nufft.u2nu(grid=np.zeros((4, 4, 4), np.complex128), coord=c,
                   forward=True, epsilon=1e-5)


@pmp("fft_order", [False, True])
@pmp("precision,vdt", [("fp64", np.complex128), ("fp32", np.complex64),
                       ("fp32", np.complex128)])
def test_fused_crop_strided_device_output(gpu, fft_order, precision, vdt):
    """2-D type 1 on a grid whose column FFT runs as two passes (the second
    one writes the uniform array itself): strided device output, both mode
    orders, compute precision independent of the output type."""
    import torch
    from ducc_amd import nufft
    rng = np.random.default_rng(21)
    shape = (400, 257)
    c, ms, _ = data(rng, shape, 1000, 2 * np.pi, False)
    eps = 1e-9 if precision == "fp64" else 1e-5
    info = {}
    big = torch.zeros((shape[0], 2 * shape[1] + 3),
                      dtype=torch.complex128 if vdt == np.complex128
                      else torch.complex64, device="cuda")
    out = big[:, 1:2 * shape[1] + 1:2]  # column stride 2
    nufft.nu2u(points=torch.from_numpy(ms.astype(vdt)).cuda(),
               coord=torch.from_numpy(c).cuda(), forward=True, epsilon=eps,
               out=out, fft_order=fft_order, precision=precision, info=info)
    # columns longer than 512 (FP32) / 256 (FP64) elements: the two-pass
    # column route
    assert info["nover"][0] > 512
    ref = O.nudft_type1(c, ms, shape, True, fft_order=fft_order)
    assert O.l2error(out.cpu().numpy(), ref) <= 2 * eps
    # untouched in between the strided columns
    assert torch.count_nonzero(big[:, 0::2]).item() == 0
