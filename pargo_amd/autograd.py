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
    focal._check_dem(dem)
    grid = focal.grid_of(dem, cellsize, nodata)
    c = _lib.default_context(dem.device.index or 0)
    c.bind_stream()
    if tuple(grad_slope.shape) != tuple(dem.shape):
        raise ValueError("grad_slope has shape %s, the DEM %s" % (tuple(grad_slope.shape), tuple(dem.shape)))
    pitch = grid.pitch

    def same_layout(t):
        return (t.dtype == torch.float32 and t.device == dem.device and t.dim() == 2
                and (t.shape[1] <= 1 or t.stride(1) == 1) and (t.shape[0] <= 1 or t.stride(0) == pitch))

    g = grad_slope
    if not same_layout(g):
        g = focal._alloc_like(dem, torch.float32)
        g.copy_(grad_slope)
    if out is None:
        out = focal._alloc_like(dem, torch.float32)
    elif tuple(out.shape) != tuple(dem.shape) or not same_layout(out):
        raise ValueError("out must be a float32 tensor of the DEM's shape and row pitch")
    c.check(_lib.lib().pargo_slope_adjoint(c.h, ctypes.byref(grid), _lib._ptr(dem), _lib._ptr(g), _lib._ptr(out)))
    return out


def slope(dem, cellsize=30.0, nodata=-9999.0):
    return SlopeFunction.apply(dem, cellsize, nodata)


class IDWFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, op):
        # device values stay on the device: no host round trip per step
        op.set_values(values.detach().to(torch.float64).contiguous())
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
