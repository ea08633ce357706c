"""Python mirror of gpro's compute-load model over the C ABI.

Three ways to obtain a compute layer, as in the reference:
  intensity ratio   Transformation(nodataLoad, validLoad, comptLayer).run()
                    (gpro/transformation.h:155-228)     -> intensity_load()
  estimate          IdwTransformation (apps/spatial/idw/idwTransformation.cpp)
                                                        -> idw.IDWOperator.compute_load()
  experiment        per-cell cost measured by a run (idwOperator.cpp:439-491)
                                                        -> idw.IDWOperator.work()
and the load-balanced row bands of ComputeLayer::getCompuLoad / valRowDcmp
(gpro/computeLayer.h:378-443, deComposition.h:305-393) -> row_bands().
Loads live on the device (float64 tensors); nothing here computes on the CPU
except the band cut, which the reference also runs serially on the host.
"""
import ctypes

from . import _lib
from .focal import grid_of
from .idw import row_bands  # noqa: F401  (re-exported)


def intensity_load(layers, nodatas, granularity, nodata_load, valid_load, work_rect=None, ctx=None):
    """Sum over the data layers (float64 device tensors of equal shape, each
    with its NoData value) of validLoad / nodataLoad per cell of the coarse
    grid of `granularity`.  work_rect = (row0, col0, rows, cols) is the data
    work rectangle (default: the interior of a 3x3 neighbourhood)."""
    import torch
    if not layers:
        raise ValueError("Datalayers used for calculating compute layer should not be null.")
    d0 = layers[0]
    rows, cols = d0.shape
    if work_rect is None:
        work_rect = (1, 1, rows - 2, cols - 2)
    r0, c0, wr, wc = (int(x) for x in work_rect)
    ctx = ctx or _lib.default_context(d0.device.index or 0)
    ctx.bind_stream()
    L = _lib.lib()
    lr, lc = ctypes.c_int(), ctypes.c_int()
    g = grid_of(d0, 1.0, nodatas[0])
    ctx.check(L.pargo_intensity_load(ctx.h, ctypes.byref(g), None, r0, c0, wr, wc, int(granularity),
                                     float(nodata_load), float(valid_load), 0, None, ctypes.byref(lr),
                                     ctypes.byref(lc)))
    out = torch.empty((lr.value, lc.value), dtype=torch.float64, device=d0.device)
    for i, (d, nd) in enumerate(zip(layers, nodatas)):
        if d.dtype != torch.float64 or d.shape != d0.shape or d.stride(1) != 1:
            raise TypeError("data layers must be float64 device tensors of one shape with contiguous rows")
        g = grid_of(d, 1.0, nd)
        ctx.check(L.pargo_intensity_load(ctx.h, ctypes.byref(g), _lib._ptr(d), r0, c0, wr, wc, int(granularity),
                                         float(nodata_load), float(valid_load), 1 if i else 0, _lib._ptr(out),
                                         ctypes.byref(lr), ctypes.byref(lc)))
    return out


class Transformation:
    """Transformation(nodataLoad, validLoad): run(layers, nodatas,
    granularity) fills and returns the compute layer."""

    def __init__(self, nodata_load, valid_load):
        if nodata_load < 0 or valid_load < 0:
            raise ValueError("The load specified cannot be negative.")
        self.nodata_load, self.valid_load = float(nodata_load), float(valid_load)

    def run(self, layers, nodatas, granularity, work_rect=None, ctx=None):
        return intensity_load(layers, nodatas, granularity, self.nodata_load, self.valid_load, work_rect, ctx)
