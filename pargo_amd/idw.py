"""Python mirror of IDWOperator over the C ABI.

Reference interface: IDWOperator(float cellsize, int nbrPoints, int idw_power,
double bufferSize, double blockSize) with readSamples / creatSampleBlocks /
idwLayer / maskLayer / Run (apps/spatial/idw/idwOperator.h:35-116).  Here
the samples come as arrays in file order, creatSampleBlocks is
pargo_idw_prepare, and run() fills a device float32 grid.
"""
import ctypes

import numpy as np

from . import _lib


def grid(xs, ys, cellsize):
    """(rows, cols, minX, maxY) of the output grid (idwOperator.cpp:126-141)."""
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    ys = np.ascontiguousarray(ys, dtype=np.float64)
    r, c = ctypes.c_int(), ctypes.c_int()
    mx, my = ctypes.c_double(), ctypes.c_double()
    rc = _lib.lib().pargo_idw_grid(_lib._ptr(xs), _lib._ptr(ys), len(xs), float(cellsize), ctypes.byref(r),
                                   ctypes.byref(c), ctypes.byref(mx), ctypes.byref(my))
    if rc != 0:
        raise _lib.PargoError(rc, "pargo_idw_grid: bad arguments")
    return r.value, c.value, mx.value, my.value


def row_bands(load, nbands, nodata=-9999.0):
    """First row of each of nbands load-balanced row bands of a host load
    grid, plus the row count (DeComposition::valRowDcmp,
    gpro/deComposition.h:305-393)."""
    load = np.ascontiguousarray(load, dtype=np.float64)
    first = np.empty(int(nbands) + 1, dtype=np.int32)
    _lib.check_host(_lib.lib().pargo_load_row_bands(_lib._ptr(load), load.shape[0], load.shape[1], float(nodata),
                                                    int(nbands), _lib._ptr(first)), "pargo_load_row_bands")
    return first


class IDWOperator:
    def __init__(self, cellsize, nbr_points, idw_power, buffer_size, block_size, ctx=None, device=0):
        self.cellsize = float(np.float32(cellsize))
        self.k = int(nbr_points)
        self.power = int(idw_power)
        self.buffer = float(buffer_size)
        self.block = float(block_size)
        self.ctx = ctx or _lib.default_context(device)
        self.plan = None
        self.n = 0
        self.mask = None
        self.mask_nodata = 0

    def read_samples(self, xs, ys, vals):
        """readSamples + creatSampleBlocks (idwOperator.cpp:63-169)."""
        self.free()
        xs = np.ascontiguousarray(xs, dtype=np.float64)
        ys = np.ascontiguousarray(ys, dtype=np.float64)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        self.n = len(xs)
        self.rows, self.cols, self.min_x, self.max_y = grid(xs, ys, self.cellsize)
        h = ctypes.c_void_p()
        self.ctx.bind_stream()
        self.ctx.check(_lib.lib().pargo_idw_prepare(self.ctx.h, _lib._ptr(xs), _lib._ptr(ys), _lib._ptr(vals),
                                                    self.n, self.cellsize, self.block, ctypes.byref(h)))
        self.plan = h
        return self

    def set_values(self, vals):
        """New sample values in file order: a host array, or a float64
        device tensor (then gathered on the device, stream-ordered, with no
        host synchronisation)."""
        import torch
        if isinstance(vals, torch.Tensor) and vals.device.type == "cuda":
            if vals.dtype != torch.float64 or vals.dim() != 1 or vals.numel() != self.n:
                raise ValueError("device values must be a float64 vector of %d samples" % self.n)
            vals = vals.contiguous()
            self.ctx.bind_stream()
            self.ctx.check(_lib.lib().pargo_idw_set_values_device(self.ctx.h, self.plan, _lib._ptr(vals)))
            return
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        if vals.shape != (self.n,):
            raise ValueError("values must hold %d samples" % self.n)
        self.ctx.check(_lib.lib().pargo_idw_set_values(self.ctx.h, self.plan, _lib._ptr(vals)))

    def mask_layer(self, mask, mask_nodata):
        """maskLayer: a device int32 grid; 0 and mask_nodata cells give -9999."""
        import torch
        _lib.check_raster(mask, torch.int32, "mask")
        self.mask = mask.contiguous()
        self.mask_nodata = int(mask_nodata)

    def _grid_tensor(self, t, dtype, name):
        """A caller-supplied (rows, cols) device tensor of dtype, contiguous
        (the kernels address it with pitch = cols)."""
        import torch
        if t is None:
            return torch.empty((self.rows, self.cols), dtype=dtype, device=torch.device("cuda", self.ctx.device))
        _lib.check_raster(t, dtype, name)
        if tuple(t.shape) != (self.rows, self.cols) or not t.is_contiguous():
            raise ValueError("%s must be a contiguous %dx%d tensor" % (name, self.rows, self.cols))
        return t

    def run(self, out=None, count_candidates=False):
        import torch
        dev = torch.device("cuda", self.ctx.device)
        out = self._grid_tensor(out, torch.float32, "out")
        cand = torch.zeros(1, dtype=torch.int64, device=dev) if count_candidates else None
        mr = mc = 0
        if self.mask is not None:
            mr, mc = self.mask.shape
        self.ctx.bind_stream()
        self.ctx.check(_lib.lib().pargo_idw_run(self.ctx.h, self.plan, self.k, self.power, self.buffer,
                                                _lib._ptr(self.mask), mr, mc, self.mask_nodata, _lib._ptr(out),
                                                _lib._ptr(cand)))
        if count_candidates:
            return out, cand
        return out

    def work(self, out=None):
        """Measured work per cell: the number of samples each cell's search
        examined (masked cells 0), float64 -- the device form of the
        reference's per-cell timing layer (-dcmp space -writeLoad,
        idwOperator.cpp:439-491).  out (float32, optional) gets the IDW
        values of the same run."""
        import torch
        if out is not None:
            out = self._grid_tensor(out, torch.float32, "out")
        work = self._grid_tensor(None, torch.float64, "work")
        mr = mc = 0
        if self.mask is not None:
            mr, mc = self.mask.shape
        self.ctx.bind_stream()
        self.ctx.check(_lib.lib().pargo_idw_work(self.ctx.h, self.plan, self.k, self.power, self.buffer,
                                                 _lib._ptr(self.mask), mr, mc, self.mask_nodata, _lib._ptr(out),
                                                 _lib._ptr(work)))
        return work

    def compute_load(self, granularity=10, mask=None, mask_nodata=0.0, nodata=-9999.0):
        """Estimated load per granularity x granularity block of cells
        (IdwTransformation over a ComputeLayer, idwTransformation.cpp:3-62),
        a float64 device tensor.  mask is a device int32 grid compared with
        mask_nodata as a double, as the reference's estimate does."""
        import torch
        L = _lib.lib()
        r, c = ctypes.c_int(), ctypes.c_int()
        self.ctx.check(L.pargo_idw_load(self.ctx.h, self.plan, self.k, int(granularity), None, 0, 0, 0.0,
                                        float(nodata), None, ctypes.byref(r), ctypes.byref(c)))
        out = torch.empty((r.value, c.value), dtype=torch.float64, device=torch.device("cuda", self.ctx.device))
        mr = mc = 0
        if mask is not None:
            _lib.check_raster(mask, torch.int32, "mask")
            mask = mask.contiguous()
            mr, mc = mask.shape
        self.ctx.bind_stream()
        self.ctx.check(L.pargo_idw_load(self.ctx.h, self.plan, self.k, int(granularity), _lib._ptr(mask), mr, mc,
                                        float(mask_nodata), float(nodata), _lib._ptr(out), ctypes.byref(r),
                                        ctypes.byref(c)))
        return out

    ORDER = {"rows": 0, "estimate": 1, "measured": 2}

    def order_tiles(self, source="estimate"):
        """Launch order of the 16 x 16 cell tiles of every later run, work and
        adjoint call: "rows" (row-major), "estimate" (heaviest first by the
        reference's load estimate, IdwTransformation over a ComputeLayer of
        granularity 16, idwTransformation.cpp:3-62) or "measured" (heaviest
        first by the samples each tile's searches examined in one pass, the
        reference's preliminary experiment, idwOperator.cpp:439-491).  The
        values do not depend on the order."""
        mr = mc = 0
        if self.mask is not None:
            mr, mc = self.mask.shape
        self.ctx.bind_stream()
        self.ctx.check(_lib.lib().pargo_idw_order_tiles(self.ctx.h, self.plan, self.k, self.power, self.buffer,
                                                        _lib._ptr(self.mask), mr, mc, self.mask_nodata,
                                                        self.ORDER[source]))

    def adjoint(self, grad_out):
        """dL/d(sample values) in file order (float64 device tensor)."""
        import torch
        grad_out = self._grid_tensor(grad_out.to(torch.float32).contiguous(), torch.float32, "grad_out")
        g = torch.empty(self.n, dtype=torch.float64, device=grad_out.device)
        mr = mc = 0
        if self.mask is not None:
            mr, mc = self.mask.shape
        self.ctx.bind_stream()
        self.ctx.check(_lib.lib().pargo_idw_adjoint(self.ctx.h, self.plan, self.k, self.power, self.buffer,
                                                    _lib._ptr(self.mask), mr, mc, self.mask_nodata,
                                                    _lib._ptr(grad_out), _lib._ptr(g)))
        return g

    def free(self):
        if self.plan:
            _lib.lib().pargo_idw_plan_free(self.ctx.h, self.plan)
            self.plan = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
