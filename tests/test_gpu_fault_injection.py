"""Allocation-failure injection (PARGO_FAIL_ALLOC) through the device entry
points and the command lines: for n = 1, 2, ... the n-th allocation point
of a call fails.  Each failure must come back as PARGO_E_NOMEM with a message
naming the point (the CLIs: a non-zero exit with an [ERROR] line, never a
signal), leave the context usable, and the call must give the reference
result once n passes its allocation points.  The reference's own failure
points are the GDAL opens and reads it checks (gpro/rasterLayer.h:752-755,
:1026-1029) and the per-cell Operator returning false
(gpro/rasterOperator.h:180-183); it stops the program there."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from pargo_amd import _lib, batch, focal, synth
from pargo_amd.idw import IDWOperator

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "pargo_amd", "_build", "bin")
DATA = os.path.join(ROOT, "tests", "golden", "data")
_gen = [0]


def _arm(n):
    _gen[0] += 1
    os.environ["PARGO_FAIL_ALLOC"] = "%d:%d" % (n, _gen[0])


def _disarm():
    os.environ.pop("PARGO_FAIL_ALLOC", None)


def _walk(call, message, max_points=16):
    """call() under the 1st, 2nd, ... failing allocation point until it
    returns PARGO_OK; returns the failed sites."""
    sites = []
    try:
        for n in range(1, max_points + 1):
            _arm(n)
            rc = call()
            _disarm()
            if rc == _lib.PARGO_OK:
                return sites
            msg = message()
            assert rc == _lib.PARGO_E_NOMEM, (n, rc, msg)
            assert "injected at" in msg, msg
            site = msg.split("injected at ")[1]
            sites.append(site[:-1] if site.endswith(")") else site)
        pytest.fail("still failing after %d points: %s" % (max_points, sites))
    finally:
        _disarm()


def test_create_alloc_and_pinned():
    L = _lib.lib()
    h = ctypes.c_void_p()
    sites = _walk(lambda: L.pargo_create(0, ctypes.byref(h)), lambda: L.pargo_last_error(None).decode())
    assert sites == ["context"]
    ctx = h
    p = ctypes.c_void_p()
    msg = lambda: L.pargo_last_error(ctx).decode()
    assert _walk(lambda: L.pargo_alloc(ctx, 4096, ctypes.byref(p)), msg) == ["pargo_alloc"]
    assert L.pargo_free(ctx, p) == 0
    assert _walk(lambda: L.pargo_alloc_host(ctx, 4096, ctypes.byref(p)), msg) == ["pargo_alloc_host"]
    assert L.pargo_free_host(ctx, p) == 0
    assert L.pargo_destroy(ctx) == 0


def test_focal_fixup_queue_and_batch():
    import torch
    z = synth.workload_dem("slope_d8", 300, 420)
    dem = torch.from_numpy(z).cuda()
    ref_s, ref_c, _ = focal.slope_d8(dem, 30.0, -9999.0)
    ctx = _lib.Context(0)
    ctx.bind_stream()
    L = _lib.lib()
    s = torch.empty_like(dem)
    c = torch.empty(dem.shape, dtype=torch.uint8, device="cuda")
    g = focal.grid_of(dem, 30.0, -9999.0)
    msg = lambda: L.pargo_last_error(ctx.h).decode()
    sites = _walk(lambda: L.pargo_slope_d8(ctx.h, ctypes.byref(g), _lib._ptr(dem), 0, 0, 256, _lib._ptr(s),
                                           _lib._ptr(c), None), msg)
    assert sites == ["fixup queue"]
    torch.cuda.synchronize()
    assert torch.equal(s, ref_s) and torch.equal(c, ref_c)
    dems = [torch.from_numpy(batch.tile_dem(t, 90 + 7 * t, 130 - 5 * t)).cuda() for t in range(6)]
    b = batch.Batch(dems, ctx=ctx)
    sites = _walk(lambda: L.pargo_batch_slope_d8(ctx.h, ctypes.cast(b.tiles, ctypes.c_void_p), len(dems), 0, 0, 256,
                                                 batch.GRAPH), msg)
    assert sites == ["batch plan (host)", "batch plan"], sites
    torch.cuda.synchronize()
    for i, d in enumerate(dems):
        rs, rc_, _ = focal.slope_d8(d, 30.0, -9999.0)
        assert torch.equal(b.slopes[i], rs) and torch.equal(b.codes[i], rc_), "tile %d" % i
    ctx.close()


def test_idw_plan_order_and_adjoint():
    import torch
    rng = np.random.default_rng(4)
    xs, ys, vs = rng.uniform(0, 3000, 300), rng.uniform(0, 2000, 300), rng.uniform(0, 100, 300)
    ref_op = IDWOperator(30.0, 12, 2, 800.0, 500.0).read_samples(xs, ys, vs)
    ref = ref_op.run().clone()
    g = torch.rand(ref.shape, device="cuda")
    ref_grad = ref_op.adjoint(g).clone()
    ctx = _lib.Context(0)
    L = _lib.lib()
    msg = lambda: L.pargo_last_error(ctx.h).decode()
    op = IDWOperator(30.0, 12, 2, 800.0, 500.0, ctx=ctx)
    plan = ctypes.c_void_p()
    xs_, ys_, vs_ = (np.ascontiguousarray(a, dtype=np.float64) for a in (xs, ys, vs))
    sites = _walk(lambda: L.pargo_idw_prepare(ctx.h, _lib._ptr(xs_), _lib._ptr(ys_), _lib._ptr(vs_), 300,
                                              ctypes.c_float(30.0), 500.0, ctypes.byref(plan)), msg)
    assert sites == ["IDW plan (host)", "IDW samples"], sites
    op.plan, op.n = plan, 300
    op.rows, op.cols = ref_op.rows, ref_op.cols
    sites = _walk(lambda: L.pargo_idw_order_tiles(ctx.h, plan, 12, 2, 800.0, None, 0, 0, 0, 2), msg)
    assert sites == ["tile loads", "tile order (host)", "tile order"], sites
    assert torch.equal(op.run(), ref)
    gg = torch.empty(300, dtype=torch.float64, device="cuda")
    sites = _walk(lambda: L.pargo_idw_adjoint(ctx.h, plan, 12, 2, 800.0, None, 0, 0, 0, _lib._ptr(g), _lib._ptr(gg)),
                  msg)
    assert sites == ["context scratch"]
    torch.cuda.synchronize()
    assert torch.allclose(gg, ref_grad, rtol=1e-12, atol=1e-15)
    op.free()
    ref_op.free()
    ctx.close()


@pytest.mark.parametrize("app,args", [
    ("slope", ["dem.tif", "moore.nbr", "OUT"]),
    ("flowdird8", ["dem.tif", "moore.nbr", "OUT", "8"]),
])
def test_command_lines_fail_cleanly(app, args, tmp_path):
    """Each allocation point of a command line, made to fail, ends the
    program with an error status and an [ERROR] message, never a signal;
    past the last point the program writes what an undisturbed run writes."""
    exe = os.path.join(BIN, app)

    def argv(out):
        # data files live in tests/golden/data; other arguments (dirType) pass as they are
        return [exe] + [str(tmp_path / out) if a == "OUT" else (os.path.join(DATA, a) if "." in a else a)
                        for a in args]

    ok = None
    for n in range(1, 60):
        env = dict(os.environ, PARGO_FAIL_ALLOC=str(n))
        p = subprocess.run(argv("out.tif"), stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env,
                           timeout=120)
        assert p.returncode >= 0, "%s killed by signal %d at point %d:\n%s" % (app, -p.returncode, n, p.stdout)
        if p.returncode == 0:
            ok = n
            break
        assert "[ERROR]" in p.stdout, p.stdout
    assert ok is not None and ok > 1, "the program never reached an allocation point"
    ref = subprocess.run(argv("ref.tif"), stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert ref.returncode == 0, ref.stdout
    L = _lib.lib()
    vals = []
    for name in ("out.tif", "ref.tif"):
        info = _lib.RasterInfo()
        path = str(tmp_path / name).encode()
        size = os.path.getsize(path) if os.path.exists(path) else -1
        rc = L.pargo_geotiff_info(path, ctypes.byref(info))
        assert rc == 0, "%s (%d bytes, last run at point %d): %s\n%s" % (
            name, size, ok, L.pargo_last_error(None).decode(), p.stdout if name == "out.tif" else ref.stdout)
        buf = np.empty(info.rows * info.cols)
        assert L.pargo_geotiff_read(path, ctypes.byref(info), _lib._ptr(buf)) == 0
        vals.append(buf)
    assert np.array_equal(vals[0], vals[1], equal_nan=True)
