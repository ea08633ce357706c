"""Many rasters at once: pargo_batch_slope_d8 / pargo_batch_autotune.

The reference handles one raster per run and spreads it over MPI ranks by
row bands (gpro/deComposition.h:148-182); a batch hands (tile, column band,
row strip) items of many rasters to the waves of one persistent launch.
"""
import ctypes
import math

import numpy as np

from . import _lib, focal, synth

PER_TILE = 0
PERSISTENT = 1
GRAPH = 2


class PargoTile(ctypes.Structure):
    """pargo_tile (include/pargo.h)."""
    _fields_ = [("grid", _lib.PargoGrid), ("dem", ctypes.c_void_p), ("slope_out", ctypes.c_void_p),
                ("code_out", ctypes.c_void_p)]


def _tile_array(dems, slopes, codes, cellsize, nodata):
    arr = (PargoTile * len(dems))()
    sizes = list(cellsize) if isinstance(cellsize, (list, tuple)) else [cellsize] * len(dems)
    if len(sizes) != len(dems):
        raise ValueError("one cell size per tile, or one for all")
    for i, d in enumerate(dems):
        arr[i].grid = focal.grid_of(d, sizes[i], nodata)
        arr[i].dem = d.data_ptr()
        arr[i].slope_out = slopes[i].data_ptr() if slopes is not None else None
        arr[i].code_out = codes[i].data_ptr() if codes is not None else None
    return arr


def _setup_lib():
    return _lib.lib()


class Batch:
    """A fixed set of device rasters and their outputs.  cellsize is one
    value for every tile or a list with one per tile."""

    def __init__(self, dems, cellsize=30.0, nodata=-9999.0, want_slope=True, want_d8=True, method="FD", unit="tan",
                 dirtype=256, ctx=None):
        import torch
        for i, d in enumerate(dems):
            _lib.check_raster(d, torch.float32, "dems[%d]" % i)
        self.dems = dems
        self.ctx = ctx or _lib.default_context(dems[0].device.index or 0)
        # outputs share each DEM's row pitch (the kernels address all three
        # rasters of a tile with one pitch)
        self.slopes = [focal._alloc_like(d, torch.float32) for d in dems] if want_slope else None
        self.codes = [focal._alloc_like(d, torch.uint8) for d in dems] if want_d8 else None
        self.tiles = _tile_array(dems, self.slopes, self.codes, cellsize, nodata)
        self.method = method if isinstance(method, int) else focal.get_slope_algorithm(method)
        self.unit = focal._unit(unit)
        self.dirtype = int(dirtype)
        self.L = _setup_lib()

    def run(self, mode=GRAPH):
        self.ctx.bind_stream()
        self.ctx.check(self.L.pargo_batch_slope_d8(self.ctx.h, ctypes.cast(self.tiles, ctypes.c_void_p),
                                                   len(self.dems), self.method, self.unit, self.dirtype, int(mode)))
        return self.slopes, self.codes

    ORDER = {"rows": 0, "estimate": 1, "measured": 2}

    def order(self, source="estimate"):
        """Queue order of later runs: heaviest tile first by cell count
        ("estimate"), by per-tile kernel time measured in one pass
        ("measured"), or as given ("rows")."""
        self.ctx.bind_stream()
        self.ctx.check(self.L.pargo_batch_order(self.ctx.h, ctypes.cast(self.tiles, ctypes.c_void_p), len(self.dems),
                                                self.method, self.unit, self.dirtype, self.ORDER[source]))

    def autotune(self):
        self.ctx.bind_stream()
        c, r = ctypes.c_int(), ctypes.c_int()
        self.ctx.check(self.L.pargo_batch_autotune(self.ctx.h, ctypes.cast(self.tiles, ctypes.c_void_p),
                                                   len(self.dems), self.method, self.unit, self.dirtype,
                                                   ctypes.byref(c), ctypes.byref(r)))
        return c.value, r.value


def tile_shapes(n, seed, side_min=256, side_max=4096):
    """Tile sizes of the batch workload: side exp(U[ln min, ln max]) and
    aspect ratio U[0.5, 2] (rows/cols), seeded."""
    rng = np.random.default_rng(seed)
    shapes = []
    for _ in range(n):
        side = math.exp(rng.uniform(math.log(side_min), math.log(side_max)))
        a = rng.uniform(0.5, 2.0)
        rows = max(3, int(round(side * math.sqrt(a))))
        cols = max(3, int(round(side / math.sqrt(a))))
        shapes.append((rows, cols))
    return shapes


def tile_dem(t, rows, cols):
    """Tile t of the batch workload: seed base+1000+t, 1-5% NoData."""
    seed = synth.BASE_SEED + 1000 + t
    z = synth.dem(seed, rows, cols)
    frac = 0.01 + 0.04 * ((seed * 2654435761) % 1000) / 1000.0
    synth.add_nodata_discs(seed, z, fraction=frac)
    return z
