"""ctypes wrapper of oracle/_build/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only to check or time the CPU restatement of the reference.
The product (pargo_amd/) never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

METHODS = {"FD": 0, "FFD": 1, "MD": 2, "SD": 3, "SFD": 4, "TFD": 5, "TFDW": 6}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("oracle not built: make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.oracle_slope.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_int, P,
                                   ctypes.c_int]
        L.oracle_d8.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, P, ctypes.c_int]
        L.oracle_idw_dims.argtypes = [P, P, ctypes.c_int, ctypes.c_float, ctypes.POINTER(ctypes.c_int),
                                      ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double)]
        L.oracle_idw.argtypes = [P, P, P, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_double, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, P, ctypes.c_int]
        L.oracle_idw_load.argtypes = [P, P, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_int, P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                      P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.oracle_val_row_dcmp.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, P]
        L.oracle_sca_d8.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, P,
                                    ctypes.POINTER(ctypes.c_int)]
        L.oracle_pit_remove.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, P, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_int)]
        L.oracle_intensity_load.argtypes = [P, ctypes.c_int, P] + [ctypes.c_int] * 7 + [ctypes.c_double,
                                                                                         ctypes.c_double, P]
        L.oracle_div_sqrt2.argtypes = [P, P, ctypes.c_int64]
        L.oracle_smpl1d_dcmp.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_int)] * 2
        L.oracle_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _widen(dem):
    # float32 -> float64 exactly (GDAL RasterIO of a Float32 band into double)
    return np.ascontiguousarray(dem, dtype=np.float32).astype(np.float64)


def slope(dem, cellsize, nodata=-9999.0, method="FD", nthreads=1):
    z = _widen(dem) if np.asarray(dem).dtype == np.float32 else np.ascontiguousarray(dem, dtype=np.float64)
    rows, cols = z.shape
    out = np.empty_like(z)
    m = METHODS[method] if isinstance(method, str) else int(method)
    rc = lib().oracle_slope(_p(z), rows, cols, float(np.float32(cellsize)), float(nodata), m, _p(out), int(nthreads))
    if rc != 0:
        raise ValueError("oracle_slope failed: %d" % rc)
    return out


def d8(dem, nodata=-9999.0, dirtype=256, nthreads=1):
    z = _widen(dem) if np.asarray(dem).dtype == np.float32 else np.ascontiguousarray(dem, dtype=np.float64)
    rows, cols = z.shape
    out = np.empty_like(z)
    rc = lib().oracle_d8(_p(z), rows, cols, float(nodata), int(dirtype), _p(out), int(nthreads))
    if rc != 0:
        raise ValueError("oracle_d8 failed: %d" % rc)
    return out


def idw_dims(xs, ys, cellsize):
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    ys = np.ascontiguousarray(ys, dtype=np.float64)
    r, c = ctypes.c_int(), ctypes.c_int()
    mnx, mxy = ctypes.c_double(), ctypes.c_double()
    rc = lib().oracle_idw_dims(_p(xs), _p(ys), len(xs), float(cellsize), ctypes.byref(r), ctypes.byref(c),
                               ctypes.byref(mnx), ctypes.byref(mxy))
    if rc != 0:
        raise ValueError("oracle_idw_dims failed")
    return r.value, c.value, mnx.value, mxy.value


def idw(xs, ys, vals, cellsize, k=12, power=2, buffer=0.0, block_size=5000.0, mask=None, mask_nodata=0,
        nthreads=1, stats=False):
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    ys = np.ascontiguousarray(ys, dtype=np.float64)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    rows, cols, _, _ = idw_dims(xs, ys, cellsize)
    out = np.empty((rows, cols), dtype=np.float64)
    cand = np.zeros(rows, dtype=np.int64)
    blk = np.zeros(rows, dtype=np.int64)
    m = None
    mr = mc = 0
    if mask is not None:
        m = np.ascontiguousarray(mask, dtype=np.int32)
        mr, mc = m.shape
    rc = lib().oracle_idw(_p(xs), _p(ys), _p(vals), len(xs), float(cellsize), int(k), int(power), float(buffer),
                          float(block_size), _p(m), mr, mc, int(mask_nodata), _p(out), _p(cand), _p(blk),
                          int(nthreads))
    if rc != 0:
        raise ValueError("oracle_idw failed: %d" % rc)
    if stats:
        return out, cand, blk
    return out


def idw_load(xs, ys, cellsize, k=12, block_size=5000.0, granularity=10, mask=None, mask_nodata=0.0,
             nodata=-9999.0):
    """IDW compute-load grid (IdwTransformation over a ComputeLayer)."""
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    ys = np.ascontiguousarray(ys, dtype=np.float64)
    r, c = ctypes.c_int(), ctypes.c_int()
    m = None
    mr = mc = 0
    if mask is not None:
        m = np.ascontiguousarray(mask, dtype=np.int32)
        mr, mc = m.shape
    L = lib()
    args = [_p(xs), _p(ys), len(xs), float(cellsize), int(k), float(block_size), int(granularity), _p(m), mr, mc,
            float(mask_nodata), float(nodata)]
    rc = L.oracle_idw_load(*args, None, ctypes.byref(r), ctypes.byref(c))
    if rc != 0:
        raise ValueError("oracle_idw_load failed: %d" % rc)
    out = np.empty((r.value, c.value), dtype=np.float64)
    rc = L.oracle_idw_load(*args, _p(out), ctypes.byref(r), ctypes.byref(c))
    if rc != 0:
        raise ValueError("oracle_idw_load failed: %d" % rc)
    return out


def val_row_dcmp(load, nodata, nsub):
    """First row of each of nsub load-balanced bands, plus rows at the end."""
    load = np.ascontiguousarray(load, dtype=np.float64)
    first = np.empty(nsub + 1, dtype=np.int32)
    rc = lib().oracle_val_row_dcmp(_p(load), load.shape[0], load.shape[1], float(nodata), int(nsub), _p(first))
    if rc != 0:
        raise ValueError("oracle_val_row_dcmp failed: %d" % rc)
    return first


def sca_d8(d8, nodata=-9999.0, converge=False):
    """Contributing area of a float64 D8 layer (window-index codes); returns
    (sca float64, number of Work sweeps)."""
    d = np.ascontiguousarray(d8, dtype=np.float64)
    rows, cols = d.shape
    out = np.empty_like(d)
    sw = ctypes.c_int(0)
    rc = lib().oracle_sca_d8(_p(d), rows, cols, float(nodata), int(bool(converge)), _p(out), ctypes.byref(sw))
    if rc != 0:
        raise RuntimeError("oracle_sca_d8 failed")
    return out, sw.value


def pit_remove(dem, nodata=-9999.0, max_sweeps=0):
    """Depression filling; returns (filled float64, sweeps, converged)."""
    z = _widen(dem) if np.asarray(dem).dtype == np.float32 else np.ascontiguousarray(dem, dtype=np.float64)
    rows, cols = z.shape
    out = np.empty_like(z)
    sw = ctypes.c_int(0)
    rc = lib().oracle_pit_remove(_p(z), rows, cols, float(nodata), _p(out), int(max_sweeps), ctypes.byref(sw))
    if rc < 0:
        raise RuntimeError("oracle_pit_remove failed")
    return out, sw.value, rc == 0


def intensity_load(layers, nodatas, row0, col0, work_rows, work_cols, g, nodata_load, valid_load):
    """Intensity-ratio load of several equally sized float64 layers."""
    a = np.ascontiguousarray(np.stack([np.asarray(l, dtype=np.float64) for l in layers]))
    n, rows, cols = a.shape
    nd = np.ascontiguousarray(nodatas, dtype=np.float64)
    lr = work_rows // g + (1 if work_rows % g else 0)
    lc = work_cols // g + (1 if work_cols % g else 0)
    out = np.empty((lr, lc))
    rc = lib().oracle_intensity_load(_p(a), n, _p(nd), rows, cols, row0, col0, work_rows, work_cols, g,
                                     float(nodata_load), float(valid_load), _p(out))
    if rc != 0:
        raise ValueError("The load specified cannot be negative.")
    return out


def div_sqrt2(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    q = np.empty_like(x)
    lib().oracle_div_sqrt2(_p(x), _p(q), x.size)
    return q


def smpl1d_dcmp(glb_begin, glb_end, nsub, isub):
    a, b = ctypes.c_int(), ctypes.c_int()
    lib().oracle_smpl1d_dcmp(glb_begin, glb_end, nsub, isub, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def num_threads():
    return lib().oracle_num_threads()
