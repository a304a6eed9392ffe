"""ducc_amd.fft.c2c against the reference's own c2c tests
(python/test/test_fft.py:25-37, 98-160: round trips within 6e-7 / 2e-15 in
l2 error over lengths 1-255, 1700-2047 and 137^2, n-d shapes, axis
subsets), against ducc0::c2c itself (compiled from the reference sources
into oracle/_ref, where built) and against numpy.fft (a C port of pocketfft,
the algorithm family of ducc0's FFT), itself pinned by the direct DFT sum
on small sizes in test_fft_host.py."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
pmp = pytest.mark.parametrize

shapes1D = ((10,), (127,))
shapes2D = ((128, 128), (128, 129), (1, 129), (2, 127), (3, 127), (6, 127),
            (129, 1), (127, 2), (127, 3), (127, 6))
shapes3D = ((32, 17, 39), (32, 1, 39), (2, 3, 17), (2, 8, 17), (5, 7, 5))
shapes = shapes1D + shapes2D + shapes3D
len1D = list(range(1, 256)) + list(range(1700, 2048)) + [137 * 137]
tol = {np.float32: 6e-7, np.float64: 2e-15}
ctype = {np.float32: np.complex64, np.float64: np.complex128}


def l2error(a, b):
    """ducc0.misc.l2error: ||a-b|| / max(||a||, ||b||), in FP64"""
    a = np.asarray(a).astype(np.complex128)
    b = np.asarray(b).astype(np.complex128)
    return np.sqrt(np.sum(np.abs(a - b) ** 2) /
                   max(np.sum(np.abs(a) ** 2), np.sum(np.abs(b) ** 2), 1e-300))


def fftn(a, axes=None, inorm=0, out=None):
    from ducc_amd import fft
    return fft.c2c(a, axes=axes, forward=True, inorm=inorm, out=out)


def ifftn(a, axes=None, inorm=0, out=None):
    from ducc_amd import fft
    return fft.c2c(a, axes=axes, forward=False, inorm=inorm, out=out)


@pmp("dtype", [np.float32, np.float64])
@pmp("inorm", [0, 1, 2])
def test1D(gpu, inorm, dtype):
    """test_fft.py:98-118 (c2c parts), all of its lengths, plus each forward
    transform against numpy.fft"""
    rng = np.random.default_rng(42)
    eps = tol[dtype]
    for n in len1D:
        a = rng.random(n) - 0.5 + 1j * rng.random(n) - 0.5j
        a = a.astype(ctype[dtype])
        f = fftn(a, inorm=inorm)
        assert f.dtype == ctype[dtype]
        ref = np.fft.fft(a.astype(np.complex128))
        ref *= {0: 1.0, 1: 1 / np.sqrt(n), 2: 1.0 / n}[inorm]
        assert l2error(f, ref) < 2 * eps, n
        assert l2error(a, ifftn(f, inorm=2 - inorm)) < eps, n
        ar = a.real
        assert l2error(ar, ifftn(fftn(ar, inorm=inorm), inorm=2 - inorm)) \
            < eps, n
        assert l2error(ar, fftn(ifftn(ar, inorm=inorm), inorm=2 - inorm)) \
            < eps, n
        assert l2error(fftn(ar.astype(ctype[dtype])), fftn(ar)) < eps, n
        tmp = a.copy()
        r = ifftn(fftn(tmp, out=tmp, inorm=inorm), out=tmp, inorm=2 - inorm)
        assert r is tmp
        assert l2error(tmp, a) < eps, n


@pmp("shp", shapes)
@pmp("inorm", [0, 1, 2])
def test_fftn(gpu, shp, inorm):
    """test_fft.py:137-147"""
    rng = np.random.default_rng(42)
    a = rng.random(shp) - 0.5 + 1j * rng.random(shp) - 0.5j
    assert l2error(a, ifftn(fftn(a, inorm=inorm), inorm=2 - inorm)) < 1e-15
    ref = np.fft.fftn(a) * {0: 1.0, 1: 1 / np.sqrt(a.size),
                            2: 1.0 / a.size}[inorm]
    assert l2error(fftn(a, inorm=inorm), ref) < 2e-15
    a = a.astype(np.complex64)
    assert l2error(a, ifftn(fftn(a, inorm=inorm), inorm=2 - inorm)) < 5e-7


@pmp("shp", shapes2D)
@pmp("axes", ((0,), (1,), (0, 1), (1, 0)))
@pmp("inorm", [0, 1, 2])
def test_fftn2D(gpu, shp, axes, inorm):
    """test_fft.py:150-160"""
    rng = np.random.default_rng(42)
    a = rng.random(shp) - 0.5 + 1j * rng.random(shp) - 0.5j
    assert l2error(a, ifftn(fftn(a, axes=axes, inorm=inorm), axes=axes,
                            inorm=2 - inorm)) < 1e-15
    assert l2error(fftn(a, axes=axes), np.fft.fftn(a, axes=axes)) < 2e-15
    a = a.astype(np.complex64)
    assert l2error(a, ifftn(fftn(a, axes=axes, inorm=inorm), axes=axes,
                            inorm=2 - inorm)) < 5e-7


@pmp("n,dtype", [(4096, np.float64), (8192, np.float32), (10007, np.float64),
                 (1 << 20, np.float32), (3 * 5 * 7 * 11 * 13 * 4, np.float64),
                 (1000003, np.float64), (65537, np.float32)])
def test_long_lines(gpu, n, dtype):
    """lines longer than a workgroup's LDS (global passes) and long prime
    lengths (Bluestein with a long convolution)"""
    rng = np.random.default_rng(5)
    a = (rng.random(n) - 0.5 + 1j * (rng.random(n) - 0.5)).astype(ctype[dtype])
    f = fftn(a)
    assert l2error(f, np.fft.fft(a.astype(np.complex128))) < 4 * tol[dtype]
    assert l2error(a, ifftn(f, inorm=2)) < 4 * tol[dtype]


@pmp("shape,axes", [((1000, 1000), None), ((3, 4096, 5), (1,)),
                    ((70000, 4), (0,)), ((6, 70000), (1,)),
                    ((17, 19, 23), (2, 0)), ((2, 3, 4, 5, 6), (1, 3, 4))])
def test_shapes_axes(gpu, shape, axes):
    """outer dimensions beyond one launch's line count, many axes"""
    rng = np.random.default_rng(6)
    a = rng.random(shape) - 0.5 + 1j * (rng.random(shape) - 0.5)
    ref = np.fft.fftn(a, axes=axes)
    assert l2error(fftn(a, axes=axes), ref) < 4e-15
    assert l2error(fftn(a.astype(np.complex64), axes=axes), ref) < 2e-6


def test_strides_and_axes_forms(gpu):
    rng = np.random.default_rng(7)
    big = rng.random((40, 66)) - 0.5 + 1j * (rng.random((40, 66)) - 0.5)
    a = big[::2, 1::3]  # strided host view
    ref = np.fft.fftn(a)
    assert l2error(fftn(a), ref) < 2e-15
    assert l2error(fftn(a, axes=(-1, -2)), ref) < 2e-15
    assert l2error(fftn(a, axes=[-1]), np.fft.fft(a, axis=1)) < 2e-15
    out = np.zeros((22, 20), np.complex128).T  # strided output
    r = fftn(a, out=out)
    assert r is out
    assert l2error(out, ref) < 2e-15
    # an axis named twice is transformed twice (N counted twice by inorm)
    twice = fftn(a, axes=(0, 0), inorm=2)
    r2 = np.fft.fft(np.fft.fft(a, axis=0), axis=0) / a.shape[0] ** 2
    assert l2error(twice, r2) < 2e-15


def test_errors(gpu):
    from ducc_amd import fft
    a = np.zeros((4, 5), np.complex128)
    with pytest.raises(RuntimeError, match="unsupported data type"):
        fft.c2c(np.zeros(4, np.int32))
    with pytest.raises(RuntimeError, match="bad axes argument"):
        fft.c2c(a, axes=[])
    with pytest.raises(RuntimeError, match="bad axes argument"):
        fft.c2c(a, axes=[0, 1, 0])
    with pytest.raises(ValueError, match="exceeds dimensionality"):
        fft.c2c(a, axes=[2])
    with pytest.raises(ValueError, match="invalid value for inorm"):
        fft.c2c(a, inorm=3)
    with pytest.raises(RuntimeError, match="incorrect data type"):
        fft.c2c(a, out=np.zeros((4, 5), np.complex64))
    with pytest.raises(RuntimeError, match="dimension mismatch"):
        fft.c2c(a, out=np.zeros((5, 4), np.complex128))


@pmp("dt", ["complex64", "complex128", "float32", "float64"])
def test_torch_device(gpu, dt):
    """device tensors in place on torch's stream, contiguous and transposed,
    against torch's own FFT (hipFFT/rocFFT)"""
    import torch
    from ducc_amd import fft
    g = torch.Generator(device="cuda").manual_seed(8)
    x = torch.randn((96, 130), generator=g, device="cuda",
                    dtype=getattr(torch, dt))
    cd = torch.complex64 if dt in ("complex64", "float32") else \
        torch.complex128
    tl = 2e-6 if cd == torch.complex64 else 1e-14
    for v in (x, x.t()):
        ref = torch.fft.fftn(v.to(cd)).cpu().numpy()
        got = fft.c2c(v)
        assert got.dtype == cd and got.is_cuda
        assert l2error(got.cpu().numpy(), ref) < tl
        back = fft.c2c(got, forward=False, inorm=2)
        assert l2error(back.cpu().numpy(), v.to(cd).cpu().numpy()) < tl
    y = x.to(cd).clone()
    r = fft.c2c(y, out=y, axes=(0,))
    assert r is y
    ref = torch.fft.fft(x.to(cd), dim=0).cpu().numpy()
    assert l2error(y.cpu().numpy(), ref) < tl
    # in place with a normalisation factor (the scaled copy reads and writes
    # the same elements)
    z = x.to(cd).clone()
    fft.c2c(z, out=z, forward=False, inorm=2)
    ref = torch.fft.ifftn(x.to(cd)).cpu().numpy()
    assert l2error(z.cpu().numpy(), ref) < tl


@pmp("shape,axes,dt", [
    ((8192,), None, np.complex128), ((6561,), None, np.complex64),
    ((137 * 137,), None, np.complex128), ((1999,), None, np.complex64),
    ((480, 1155), None, np.complex128), ((256, 243), (1,), np.complex64),
    ((97, 64, 50), None, np.complex128), ((60, 61, 62), (0, 2), np.complex64)])
@pmp("forward", [True, False])
def test_vs_reference_c2c(gpu, shape, axes, dt, forward):
    """same transform as ducc0::c2c compiled from the reference sources"""
    from oracle import ref
    if not ref.available():
        pytest.skip("reference library not built (needs the ducc0 sources)")
    rng = np.random.default_rng(9)
    a = (rng.random(shape) - 0.5 + 1j * (rng.random(shape) - 0.5)).astype(dt)
    r = ref.c2c(a, axes=axes, forward=forward, fct=0.5)
    from ducc_amd import fft
    g = fft.c2c(a, axes=axes, forward=forward)
    t = 2 * tol[np.float32 if dt == np.complex64 else np.float64]
    assert l2error(0.5 * g, r) < t


def test_back_to_back_device_calls(gpu):
    """calls on device tensors return once queued: their scratch and
    Bluestein buffers must outlive the kernels (checked after all calls)"""
    import torch
    from ducc_amd import fft
    g = torch.Generator(device="cuda").manual_seed(10)
    shapes = [(257, 130), (64, 1999), (3, 4096, 5), (6001,), (97, 101),
              (128, 128)]
    ins, outs = [], []
    for shp in shapes:
        x = torch.randn(shp, generator=g, device="cuda",
                        dtype=torch.complex128)
        ins.append(x)
        outs.append(fft.c2c(x, inorm=1))
    torch.cuda.synchronize()
    for x, y in zip(ins, outs):
        ref = torch.fft.fftn(x, norm="ortho").cpu().numpy()
        assert l2error(y.cpu().numpy(), ref) < 1e-14
