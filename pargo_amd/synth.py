"""Deterministic synthetic inputs (host side, numpy) from libpargo's generator.

Seeds: base 0x5041524730 ("PARG0"); workload i uses base + i and batch tile t
uses base + 1000 + t.
"""
import ctypes
import os

import numpy as np

from . import _lib

BASE_SEED = 0x5041524730
NODATA = -9999.0


def _nthreads():
    return max(1, min(16, os.cpu_count() or 1))


def dem(seed, rows, cols, nthreads=None):
    out = np.empty((rows, cols), dtype=np.float32)
    rc = _lib.lib().pargo_synth_dem(ctypes.c_uint64(seed), rows, cols, out.ctypes.data_as(ctypes.c_void_p),
                                    int(nthreads or _nthreads()))
    if rc != 0:
        raise RuntimeError("pargo_synth_dem failed: %d" % rc)
    return out


def add_plateaus(seed, z, fraction):
    rows, cols = z.shape
    rc = _lib.lib().pargo_synth_plateaus(ctypes.c_uint64(seed), rows, cols, z.ctypes.data_as(ctypes.c_void_p),
                                         float(fraction))
    if rc != 0:
        raise RuntimeError("pargo_synth_plateaus failed: %d" % rc)
    return z


def add_nodata_discs(seed, z, n_discs=0, fraction=0.0, nodata=NODATA):
    rows, cols = z.shape
    cov = ctypes.c_int64()
    rc = _lib.lib().pargo_synth_nodata_discs(ctypes.c_uint64(seed), rows, cols, z.ctypes.data_as(ctypes.c_void_p),
                                             int(n_discs), float(fraction), float(nodata), ctypes.byref(cov))
    if rc != 0:
        raise RuntimeError("pargo_synth_nodata_discs failed: %d" % rc)
    return cov.value


def workload_dem(kind, rows=None, cols=None, seed=None):
    """The benchmark DEMs.
    "slope1k":  1024^2, 20 NoData discs (the reference-runnable slope case).
    "slope_d8": 16384^2, 2% plateaus then 1% NoData (slope+D8 benchmark).
    Smaller rows/cols give the same recipe at another size."""
    if kind == "slope1k":
        rows = rows or 1024
        cols = cols or 1024
        seed = BASE_SEED + 0 if seed is None else seed
        z = dem(seed, rows, cols)
        add_nodata_discs(seed, z, n_discs=20)
        return z
    if kind == "slope_d8":
        rows = rows or 16384
        cols = cols or 16384
        seed = BASE_SEED + 1 if seed is None else seed
        z = dem(seed, rows, cols)
        add_plateaus(seed, z, 0.02)
        add_nodata_discs(seed, z, fraction=0.01)
        return z
    raise ValueError(kind)


def samples(seed, n, x0=0.0, y0=0.0, width=122880.0, height=122880.0, cellsize=30.0):
    xs = np.empty(n, dtype=np.float64)
    ys = np.empty(n, dtype=np.float64)
    vs = np.empty(n, dtype=np.float64)
    P = ctypes.c_void_p
    rc = _lib.lib().pargo_synth_samples(ctypes.c_uint64(seed), n, x0, y0, width, height, cellsize,
                                        xs.ctypes.data_as(P), ys.ctypes.data_as(P), vs.ctypes.data_as(P))
    if rc != 0:
        raise RuntimeError("pargo_synth_samples failed: %d" % rc)
    return xs, ys, vs
