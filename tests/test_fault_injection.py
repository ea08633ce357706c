"""Allocation-failure injection (PARGO_FAIL_ALLOC) through the host-only C
ABI entry points (CPU only): for n = 1, 2, ... the n-th allocation point of
one call fails.  Every such call must return PARGO_E_NOMEM with a message
naming the point, leave no partial result that a later call sees, and the
same call must succeed once n passes the call's allocation points.  The
reference stops with a message at the same kinds of points (failed opens
and reads, gpro/rasterLayer.h:752-755, :1026-1029; neighbourhood file
errors, gpro/neighborhood.h:250-254)."""
import ctypes
import os

import numpy as np
import pytest

from pargo_amd import _lib

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "data")
_gen = [0]


def _arm(n):
    """Selects the n-th allocation point from now (a fresh value each time)."""
    _gen[0] += 1
    os.environ["PARGO_FAIL_ALLOC"] = "%d:%d" % (n, _gen[0])


def _disarm():
    os.environ.pop("PARGO_FAIL_ALLOC", None)


def _walk(call, check_ok, max_points=12):
    """Runs call() with the 1st, 2nd, ... allocation point failing until it
    succeeds; returns the sites that failed."""
    L = _lib.lib()
    sites = []
    try:
        for n in range(1, max_points + 1):
            _arm(n)
            rc = call()
            if rc == _lib.PARGO_OK:
                _disarm()
                check_ok()
                return sites
            msg = L.pargo_last_error(None).decode()
            assert rc == _lib.PARGO_E_NOMEM, (n, rc, msg)
            assert "injected at" in msg, msg
            site = msg.split("injected at ")[1]
            sites.append(site[:-1] if site.endswith(")") else site)
        pytest.fail("still failing after %d allocation points: %s" % (max_points, sites))
    finally:
        _disarm()


def test_geotiff_read_and_info():
    L = _lib.lib()
    info = _lib.RasterInfo()
    path = os.path.join(DATA, "dem.tif").encode()
    assert L.pargo_geotiff_info(path, ctypes.byref(info)) == 0
    rows, cols = info.rows, info.cols
    buf = np.full(rows * cols, np.nan)
    ref = np.empty(rows * cols)
    assert L.pargo_geotiff_read(path, ctypes.byref(info), _lib._ptr(ref)) == 0

    def ok():
        assert np.array_equal(buf, ref)

    sites = _walk(lambda: L.pargo_geotiff_read(path, ctypes.byref(info), _lib._ptr(buf)), ok)
    assert sites == ["geotiff.file", "geotiff.raster", "geotiff.block"], sites
    assert _walk(lambda: L.pargo_geotiff_info(path, ctypes.byref(info)), lambda: None) == ["geotiff.file"]


def test_geotiff_write(tmp_path):
    L = _lib.lib()
    info = _lib.RasterInfo()
    info.rows, info.cols = 7, 9
    info.transform[:] = [100.0, 30.0, 0.0, 500.0, 0.0, -30.0]
    info.has_nodata, info.nodata = 1, -9999.0
    data = np.arange(63, dtype=np.float64)
    out = str(tmp_path / "w.tif").encode()

    def ok():
        back = np.empty(63)
        assert L.pargo_geotiff_read(out, ctypes.byref(_lib.RasterInfo()), _lib._ptr(back)) == 0
        assert np.array_equal(back, data)

    sites = _walk(lambda: L.pargo_geotiff_write(out, ctypes.byref(info), _lib._ptr(data), 0, None), ok)
    assert sites == ["geotiff.header", "geotiff.row"], sites


def test_points_read():
    L = _lib.lib()
    shp = os.path.join(DATA, "samples.shp").encode()
    n = ctypes.c_int(0)
    assert L.pargo_points_read(shp, 0, ctypes.byref(n), None, None, None) == 0
    xs, ys, vs = np.empty(n.value), np.empty(n.value), np.empty(n.value)

    def ok():
        assert np.all(np.isfinite(xs)) and np.all(np.isfinite(vs))

    m = ctypes.c_int(n.value)
    sites = _walk(lambda: L.pargo_points_read(shp, 0, ctypes.byref(m), _lib._ptr(xs), _lib._ptr(ys), _lib._ptr(vs)),
                  ok)
    assert sites == ["shapefile.file", "shapefile.file", "shapefile.file", "shapefile.points"], sites


def test_load_row_bands():
    L = _lib.lib()
    load = np.random.default_rng(3).uniform(0, 5, (40, 6))
    first = np.zeros(5, dtype=np.int32)
    ref = np.zeros(5, dtype=np.int32)
    assert L.pargo_load_row_bands(_lib._ptr(load), 40, 6, -9999.0, 4, _lib._ptr(ref)) == 0

    def ok():
        assert np.array_equal(first, ref)

    sites = _walk(lambda: L.pargo_load_row_bands(_lib._ptr(load), 40, 6, -9999.0, 4, _lib._ptr(first)), ok)
    assert sites == ["bands.row_loads"]


def test_unset_or_zero_injects_nothing():
    L = _lib.lib()
    load = np.ones((10, 3))
    first = np.zeros(3, dtype=np.int32)
    for v in ("", "0", "-3", "junk"):
        os.environ["PARGO_FAIL_ALLOC"] = v
        try:
            assert L.pargo_load_row_bands(_lib._ptr(load), 10, 3, -9999.0, 2, _lib._ptr(first)) == 0
        finally:
            _disarm()
