"""pargo_slope_d8_noalloc (the entry point that allocates nothing) against
pargo_slope_d8 through the C ABI: identical slope, D8 and NoData-mask bytes
on two calls in a row, the second on a smaller raster, each with special
values (Inf, NaN, |z| >= 2^40) that send strips through the fixup queue the
static scratch array holds."""
import ctypes

import numpy as np
import pytest

from pargo_amd import _lib, focal, synth

pytestmark = pytest.mark.gpu


def _case(rows, cols, seed):
    z = synth.workload_dem("slope_d8", rows, cols, seed=seed)
    z[5, 5] = np.inf
    z[rows // 2, 7] = np.nan
    z[10, cols - 3] = 3.0e12
    z[rows - 4, cols // 3] = -np.inf
    return z


def _outputs(dem):
    import torch
    rows, cols = dem.shape
    return (torch.full((rows, cols), float("nan"), dtype=torch.float32, device=dem.device),
            torch.full((rows, cols), 77, dtype=torch.uint8, device=dem.device),
            torch.full((rows, (cols + 31) // 32), -1, dtype=torch.int32, device=dem.device))


def test_noalloc_equals_slope_d8_two_calls():
    import torch
    L = _lib.lib()
    # a context of its own: it has never run pargo_slope_d8, so it holds no
    # allocated fixup queue the no-alloc call could lean on
    ctx = _lib.Context(0)
    ctx.bind_stream()
    for (rows, cols), seed in (((700, 1300), 11), ((123, 517), 12)):
        print("noalloc case %dx%d seed %d" % (rows, cols, seed), flush=True)
        dem = torch.from_numpy(_case(rows, cols, seed)).cuda()
        g = focal.grid_of(dem, 30.0, -9999.0)
        s0, c0, m0 = _outputs(dem)
        s1, c1, m1 = _outputs(dem)
        ctx.check(L.pargo_slope_d8_noalloc(ctx.h, ctypes.byref(g), _lib._ptr(dem), 0, 0, 256, _lib._ptr(s1),
                                           _lib._ptr(c1), _lib._ptr(m1)))
        ref = _lib.default_context(0)
        ref.bind_stream()
        ref.check(L.pargo_slope_d8(ref.h, ctypes.byref(g), _lib._ptr(dem), 0, 0, 256, _lib._ptr(s0), _lib._ptr(c0),
                                   _lib._ptr(m0)))
        torch.cuda.synchronize()
        assert torch.equal(c1, c0), "D8 codes differ (%dx%d)" % (rows, cols)
        assert torch.equal(m1, m0), "NoData masks differ (%dx%d)" % (rows, cols)
        assert torch.equal(s1.view(torch.int32), s0.view(torch.int32)), "slope bits differ (%dx%d)" % (rows, cols)
    ctx.close()


def test_noalloc_rejects_what_slope_d8_rejects():
    import torch
    L = _lib.lib()
    ctx = _lib.default_context(0)
    ctx.bind_stream()
    dem = torch.zeros((2, 40), dtype=torch.float32, device="cuda")
    out = torch.empty_like(dem)
    code = torch.empty(dem.shape, dtype=torch.uint8, device="cuda")
    g = focal.grid_of(dem, 30.0, -9999.0)
    rc = L.pargo_slope_d8_noalloc(ctx.h, ctypes.byref(g), _lib._ptr(dem), 0, 0, 256, _lib._ptr(out),
                                  _lib._ptr(code), None)
    assert rc == _lib.PARGO_E_INVALID
    dem = torch.zeros((8, 40), dtype=torch.float32, device="cuda")
    out = torch.empty_like(dem)
    code = torch.empty(dem.shape, dtype=torch.uint8, device="cuda")
    g = focal.grid_of(dem, 30.0, -9999.0)
    rc = L.pargo_slope_d8_noalloc(ctx.h, ctypes.byref(g), _lib._ptr(dem), 0, 0, 7, _lib._ptr(out),
                                  _lib._ptr(code), None)
    assert rc == _lib.PARGO_E_INVALID
    assert b"dirType" in L.pargo_last_error(ctx.h)
