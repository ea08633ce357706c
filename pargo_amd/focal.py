"""Python mirror of the focal operators (SlopeOperator, D8Operator) over the
C ABI.  Rasters are torch tensors on a HIP device; nothing here computes on
the CPU.

Reference interfaces mirrored:
  SlopeOperator{demLayer, slopeLayer, calcAlgorithm} + Run()
      apps/morphology/slope/slopeOperator.h:30-55, slopeOperator.cpp:28-44
  D8Operator(int dirType){demLayer, d8Layer} + Run()
      apps/hydrology/flowdird8/d8Operator.h:13-40, d8Operator.cpp:5-19
  GetSlopeAlgorithm (unknown names fall back to FD)
      apps/morphology/slope/slopeOperator.cpp:3-26
"""
import ctypes

from . import _lib

METHODS = _lib.SLOPE_METHODS


def get_slope_algorithm(name):
    """GetSlopeAlgorithm: exact (case-insensitive, like StringMatch) names,
    FD otherwise (slopeOperator.cpp:3-26, utility.cpp:40-47)."""
    return METHODS.get(str(name).upper(), METHODS["FD"])


def _check_dem(dem):
    import torch
    _lib.check_raster(dem, torch.float32, "dem")


def grid_of(dem, cellsize, nodata):
    g = _lib.PargoGrid()
    g.rows, g.cols = int(dem.shape[0]), int(dem.shape[1])
    g.pitch = _lib.row_pitch(dem)
    g.cellsize = float(cellsize)
    g.nodata = float(nodata)
    return g


def _unit(unit):
    if unit in ("tan", 0, None):
        return _lib.UNIT_TAN
    if unit in ("degrees", "deg", 1):
        return _lib.UNIT_DEGREES
    raise ValueError("unknown unit %r" % (unit,))


def _alloc_like(dem, dtype):
    import torch
    rows, cols = dem.shape
    pitch = _lib.row_pitch(dem)
    buf = torch.empty((rows, pitch), dtype=dtype, device=dem.device)
    return buf[:, :cols]


def _mask_buf(dem):
    import torch
    rows, cols = dem.shape
    return torch.empty((rows, (cols + 31) // 32), dtype=torch.int32, device=dem.device)


def slope_d8(dem, cellsize, nodata=-9999.0, method="FD", unit="tan", dirtype=256, want_slope=True,
             want_d8=True, want_mask=False, ctx=None):
    """Fused slope + D8 on the GPU.  Returns (slope f32, code u8, mask i32 words);
    entries not requested are None."""
    _check_dem(dem)
    ctx = ctx or _lib.default_context(dem.device.index or 0)
    ctx.bind_stream()
    import torch
    slope = _alloc_like(dem, torch.float32) if want_slope else None
    code = _alloc_like(dem, torch.uint8) if want_d8 else None
    mask = _mask_buf(dem) if want_mask else None
    g = grid_of(dem, cellsize, nodata)
    m = method if isinstance(method, int) else get_slope_algorithm(method)
    L = _lib.lib()
    if want_slope and want_d8:
        rc = L.pargo_slope_d8(ctx.h, ctypes.byref(g), _lib._ptr(dem), m, _unit(unit), int(dirtype),
                              _lib._ptr(slope), _lib._ptr(code), _lib._ptr(mask))
    elif want_slope:
        rc = L.pargo_slope(ctx.h, ctypes.byref(g), _lib._ptr(dem), m, _unit(unit), _lib._ptr(slope),
                           _lib._ptr(mask))
    else:
        rc = L.pargo_d8(ctx.h, ctypes.byref(g), _lib._ptr(dem), int(dirtype), _lib._ptr(code), _lib._ptr(mask))
    ctx.check(rc)
    return slope, code, mask


def slope(dem, cellsize, nodata=-9999.0, method="FD", unit="tan", want_mask=False, ctx=None):
    s, _, m = slope_d8(dem, cellsize, nodata, method, unit, want_d8=False, want_mask=want_mask, ctx=ctx)
    return (s, m) if want_mask else s


def d8(dem, cellsize=1.0, nodata=-9999.0, dirtype=256, want_mask=False, ctx=None):
    _, c, m = slope_d8(dem, cellsize, nodata, "FD", "tan", dirtype, want_slope=False, want_mask=want_mask, ctx=ctx)
    return (c, m) if want_mask else c


def unpack_mask(mask_words, cols):
    """NoData bitmask words -> bool tensor (rows, cols)."""
    import torch
    w = mask_words.to(torch.int64) & 0xFFFFFFFF
    bits = torch.arange(32, device=w.device, dtype=torch.int64)
    b = ((w.unsqueeze(-1) >> bits) & 1).bool()
    return b.reshape(w.shape[0], -1)[:, :cols]


def codes_to_reference(code, nodata_mask, dirtype, nodata):
    """Device codes -> the reference's double layer values (noData, -1, codes)."""
    import torch
    out = code.to(torch.float64)
    flat = PARGO_FLAT[int(dirtype)]
    out = torch.where(code == flat, torch.full_like(out, -1.0), out)
    out = torch.where(nodata_mask, torch.full_like(out, float(nodata)), out)
    return out


PARGO_FLAT = {256: _lib.D8_FLAT_256, 8: _lib.D8_FLAT_8}
