"""Python mirror of the hydrology operators over the C ABI: D8 contributing
area (SCAOperator) with its input conversion, and depression filling
(PitRemoveOperator).  Rasters are torch tensors on a HIP device; nothing here
computes on the CPU.

Reference interfaces mirrored:
  SCAOperator{d8Layer, scaLayer} + Run()
      apps/hydrology/scad8/scaOperator.h:13-41, scaOperator.cpp:5-131
  PitRemoveOperator{demLayer, wdemLayer} + Run()
      apps/hydrology/pitremove/pitRemoveOperator.h:87-113, pitRemoveOperator.cpp:3-72
"""
import ctypes

from . import _lib
from .focal import grid_of

MODES = {"reference": _lib.SCA_REFERENCE, "converge": _lib.SCA_CONVERGE}


def _check(t, dtype, name):
    _lib.check_raster(t, dtype, name)


def d8_classify(d8, nodata, ctx=None):
    """Reference D8 layer (float64 window-index codes) -> uint8 flow classes."""
    import torch
    _check(d8, torch.float64, "d8")
    ctx = ctx or _lib.default_context(d8.device.index or 0)
    ctx.bind_stream()
    rows, cols = d8.shape
    buf = torch.empty((rows, _lib.row_pitch(d8)), dtype=torch.uint8, device=d8.device)[:, :cols]
    g = grid_of(d8, 1.0, nodata)
    ctx.check(_lib.lib().pargo_d8_classify(ctx.h, ctypes.byref(g), _lib._ptr(d8), _lib._ptr(buf)))
    return buf


def sca_d8(flow, nodata, nodata_mask=None, mode="reference", ctx=None):
    """Contributing area (cells) from flow classes (uint8; pargo_d8's
    dirtype-8 codes qualify, with its NoData mask).  Returns float64 with the
    reference's NoData on the border and on NoData cells."""
    import torch
    _check(flow, torch.uint8, "flow")
    if mode not in MODES:
        raise ValueError("mode must be one of %s" % sorted(MODES))
    if nodata_mask is not None:
        _check(nodata_mask, torch.int32, "nodata_mask")
        if nodata_mask.shape != (flow.shape[0], (flow.shape[1] + 31) // 32) or not nodata_mask.is_contiguous():
            raise ValueError("nodata_mask must hold ceil(cols/32) words per row, contiguous")
    ctx = ctx or _lib.default_context(flow.device.index or 0)
    ctx.bind_stream()
    rows, cols = flow.shape
    out = torch.empty((rows, _lib.row_pitch(flow)), dtype=torch.float64, device=flow.device)[:, :cols]
    g = grid_of(flow, 1.0, nodata)
    ctx.check(_lib.lib().pargo_sca_d8(ctx.h, ctypes.byref(g), _lib._ptr(flow), _lib._ptr(nodata_mask),
                                      MODES[mode], _lib._ptr(out)))
    return out


class SCAOperator:
    """SCAOperator: d8Layer(layer) then scaLayer() / Run().  The D8 layer is
    the reference's float64 layer of dirType-8 codes with its NoData."""

    def __init__(self, mode="reference"):
        self.mode = mode
        self._d8 = None
        self._nodata = None
        self.sca = None

    def d8_layer(self, d8, nodata):
        self._d8, self._nodata = d8, float(nodata)
        return self

    def run(self, ctx=None):
        if self._d8 is None:
            raise RuntimeError("d8_layer() first")
        flow = d8_classify(self._d8, self._nodata, ctx)
        self.sca = sca_d8(flow, self._nodata, None, self.mode, ctx)
        return self.sca


def pit_remove(dem, nodata, ctx=None):
    """Depression filling (PitRemoveOperator) of a float32 DEM on the device.
    Returns (filled float64 with the reference's values, device passes)."""
    import torch
    _check(dem, torch.float32, "dem")
    ctx = ctx or _lib.default_context(dem.device.index or 0)
    ctx.bind_stream()
    rows, cols = dem.shape
    out = torch.empty((rows, _lib.row_pitch(dem)), dtype=torch.float64, device=dem.device)[:, :cols]
    g = grid_of(dem, 1.0, nodata)
    passes = ctypes.c_int(0)
    ctx.check(_lib.lib().pargo_pit_remove(ctx.h, ctypes.byref(g), _lib._ptr(dem), _lib._ptr(out),
                                          ctypes.byref(passes)))
    return out, passes.value


class PitRemoveOperator:
    """PitRemoveOperator: demLayer(layer) then Run(); the filled layer is
    returned by run() and kept in .wdem."""

    def __init__(self):
        self._dem = None
        self._nodata = None
        self.wdem = None
        self.passes = 0

    def dem_layer(self, dem, nodata):
        self._dem, self._nodata = dem, float(nodata)
        return self

    def run(self, ctx=None):
        if self._dem is None:
            raise RuntimeError("dem_layer() first")
        self.wdem, self.passes = pit_remove(self._dem, self._nodata, ctx)
        return self.wdem
