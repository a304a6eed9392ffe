"""torch.autograd wrappers of the GPU NUFFT.

u2nu (type 2): out_j = sum_k grid_k exp(s i 2pi/P k.x_j), s = -1 if forward.
  * grad w.r.t. grid: the adjoint, i.e. the type-1 transform with the
    opposite sign (nu2u of the output gradient);
  * grad w.r.t. coord: Re(g_j * conj(D_jd)) with D_jd = d out_j / d x_jd
    from the kernel-derivative interpolation (dmx_u2nu_coord_grad).
nu2u (type 1): out_k = sum_j points_j exp(s i 2pi/P k.x_j).
  * grad w.r.t. points: type 2 with the opposite sign;
  * grad w.r.t. coord: Re(points_j * conj(D'_jd)), D' the coordinate
    derivative of the opposite-sign type 2 of the output gradient.

Gradients follow torch's convention for complex tensors (grad = dL/dRe +
i dL/dIm).  Tensors must be on the GPU; coordinates float32/float64, values
complex64/complex128.
"""
import numpy as np
import torch

from . import nufft

_TWO_PI = 2 * np.pi


def _kw(ctx):
    return dict(epsilon=ctx.eps, periodicity=ctx.periodicity,
                fft_order=ctx.fft_order, precision=ctx.precision)


class U2NU(torch.autograd.Function):
    @staticmethod
    def forward(ctx, grid, coord, forward, eps, periodicity, fft_order,
                precision):
        ctx.fwd, ctx.eps, ctx.periodicity = forward, eps, periodicity
        ctx.fft_order, ctx.precision = fft_order, precision
        g = grid.detach().contiguous()
        c = coord.detach().contiguous()
        ctx.save_for_backward(g, c)
        return nufft.u2nu(grid=g, coord=c, forward=forward, **_kw(ctx))

    @staticmethod
    def backward(ctx, gout):
        g, c = ctx.saved_tensors
        gout = gout.contiguous()
        ggrid = gcoord = None
        if ctx.needs_input_grad[0]:
            ggrid = torch.empty_like(g)
            nufft.nu2u(points=gout, coord=c, forward=not ctx.fwd, out=ggrid,
                       **_kw(ctx))
        if ctx.needs_input_grad[1]:
            d = nufft.u2nu_coord_grad(grid=g, coord=c, forward=ctx.fwd,
                                      **_kw(ctx))
            gcoord = (gout[:, None] * d.conj()).real.to(c.dtype)
        return ggrid, gcoord, None, None, None, None, None


class NU2U(torch.autograd.Function):
    @staticmethod
    def forward(ctx, points, coord, shape, forward, eps, periodicity,
                fft_order, precision):
        ctx.fwd, ctx.eps, ctx.periodicity = forward, eps, periodicity
        ctx.fft_order, ctx.precision = fft_order, precision
        p = points.detach().contiguous()
        c = coord.detach().contiguous()
        ctx.save_for_backward(p, c)
        out = torch.empty(tuple(shape), dtype=p.dtype, device=p.device)
        return nufft.nu2u(points=p, coord=c, forward=forward, out=out,
                          **_kw(ctx))

    @staticmethod
    def backward(ctx, gout):
        p, c = ctx.saved_tensors
        gout = gout.contiguous()
        gp = gcoord = None
        if ctx.needs_input_grad[0]:
            gp = nufft.u2nu(grid=gout, coord=c, forward=not ctx.fwd,
                            **_kw(ctx))
        if ctx.needs_input_grad[1]:
            d = nufft.u2nu_coord_grad(grid=gout, coord=c, forward=not ctx.fwd,
                                      **_kw(ctx))
            gcoord = (p[:, None] * d.conj()).real.to(c.dtype)
        return gp, gcoord, None, None, None, None, None, None


def u2nu(grid, coord, forward=True, eps=1e-6, periodicity=_TWO_PI,
         fft_order=False, precision="auto"):
    """Differentiable type-2 NUFFT (grid -> values at coord)."""
    return U2NU.apply(grid, coord, forward, eps, periodicity, fft_order,
                      precision)


def nu2u(points, coord, shape, forward=True, eps=1e-6, periodicity=_TWO_PI,
         fft_order=False, precision="auto"):
    """Differentiable type-1 NUFFT (values at coord -> grid of `shape`)."""
    return NU2U.apply(points, coord, tuple(shape), forward, eps, periodicity,
                      fft_order, precision)
