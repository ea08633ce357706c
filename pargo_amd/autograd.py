"""torch.autograd wrappers over the C ABI's forward and adjoint kernels.

SlopeFunction: slope(z) (Horn, tan units) with dL/dz from
pargo_slope_adjoint; only the DEM is saved for the backward pass.
IDWFunction: IDW grid as a function of the sample values, with
dL/d(values) from pargo_idw_adjoint (the sample geometry is fixed).

Both run on the GPU through libpargo.so; there is no CPU fallback.
"""
import ctypes

import torch

from . import _lib, focal


class SlopeFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, dem, cellsize, nodata):
        s = focal.slope(dem, cellsize, nodata, "FD", "tan")
        ctx.save_for_backward(dem)
        ctx.cellsize = cellsize
        ctx.nodata = nodata
        return s

    @staticmethod
    def backward(ctx, grad_slope):
        (dem,) = ctx.saved_tensors
        return slope_adjoint(dem, grad_slope, ctx.cellsize, ctx.nodata), None, None


def slope_adjoint(dem, grad_slope, cellsize, nodata, out=None):
    c = _lib.default_context(dem.device.index or 0)
    c.bind_stream()
    g = grad_slope.contiguous() if grad_slope.stride(1) != 1 else grad_slope
    if g.stride(0) != dem.stride(0):
        g2 = torch.empty_strided(dem.shape, dem.stride(), dtype=torch.float32, device=dem.device)
        g2.copy_(g)
        g = g2
    if out is None:
        out = torch.empty_strided(dem.shape, dem.stride(), dtype=torch.float32, device=dem.device)
    grid = focal.grid_of(dem, cellsize, nodata)
    c.check(_lib.lib().pargo_slope_adjoint(c.h, ctypes.byref(grid), _lib._ptr(dem), _lib._ptr(g), _lib._ptr(out)))
    return out


def slope(dem, cellsize=30.0, nodata=-9999.0):
    return SlopeFunction.apply(dem, cellsize, nodata)


class IDWFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, op):
        op.set_values(values.detach().cpu().double().numpy())
        out = op.run()
        ctx.op = op
        return out

    @staticmethod
    def backward(ctx, grad_out):
        g = ctx.op.adjoint(grad_out.float())
        return g, None


def idw(values, op):
    """IDW grid of `values` (file order, a float64 tensor) with the sample
    geometry and parameters of the IDWOperator `op`."""
    return IDWFunction.apply(values, op)
