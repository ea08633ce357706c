"""GeoTIFF and point-shapefile files through libpargo's GDAL-free codecs.

Mirrors what RasterLayer::readFile / writeFile (gpro/rasterLayer.h:749-778,
882-903, 957-1087) and IDWOperator::readSamples
(apps/spatial/idw/idwOperator.cpp:63-149) take from GDAL and OGR.  All
functions run on the host; none needs a GPU.
"""
import ctypes

import numpy as np

from . import _lib

FLOAT64, FLOAT32, INT32, UINT8 = 0, 1, 2, 3
_TYPES = {"float64": FLOAT64, "float32": FLOAT32, "int32": INT32, "uint8": UINT8}


def _info_dict(ri):
    return {"rows": ri.rows, "cols": ri.cols, "transform": tuple(ri.transform),
            "nodata": ri.nodata if ri.has_nodata else None, "bits": ri.bits,
            "sample_format": ri.sample_format,
            # RasterLayer's cellSize is float(pTransform[1]) (gpro/metaData.h:56)
            "cellsize": float(np.float32(ri.transform[1]))}


def geotiff_info(path):
    ri = _lib.RasterInfo()
    _lib.check_host(_lib.lib().pargo_geotiff_info(path.encode(), ctypes.byref(ri)), "pargo_geotiff_info")
    return _info_dict(ri)


def read_geotiff(path):
    """Band 1 as a float64 (rows, cols) array and the header dict."""
    info = geotiff_info(path)
    out = np.empty((info["rows"], info["cols"]), dtype=np.float64)
    ri = _lib.RasterInfo()
    _lib.check_host(_lib.lib().pargo_geotiff_read(path.encode(), ctypes.byref(ri), _lib._ptr(out)),
                    "pargo_geotiff_read")
    return out, _info_dict(ri)


def write_geotiff(path, data, transform, nodata=None, dtype="float64", georef_from=None):
    """Writes a 2-D array as a one-band GeoTIFF of `dtype`."""
    a = np.ascontiguousarray(data, dtype=np.float64)
    if a.ndim != 2:
        raise ValueError("write_geotiff: expected a 2-D array")
    ri = _lib.RasterInfo()
    ri.rows, ri.cols = a.shape
    for i, v in enumerate(transform):
        ri.transform[i] = float(v)
    ri.has_nodata = 0 if nodata is None else 1
    ri.nodata = 0.0 if nodata is None else float(nodata)
    _lib.check_host(_lib.lib().pargo_geotiff_write(path.encode(), ctypes.byref(ri), _lib._ptr(a), _TYPES[dtype],
                                                   georef_from.encode() if georef_from else None),
                    "pargo_geotiff_write")


def read_points(shp_path, field_index):
    """x, y and DBF field `field_index` of every point (float64 arrays)."""
    L = _lib.lib()
    n = ctypes.c_int(0)
    _lib.check_host(L.pargo_points_read(shp_path.encode(), int(field_index), ctypes.byref(n), None, None, None),
                    "pargo_points_read")
    xs, ys, vs = (np.empty(n.value, dtype=np.float64) for _ in range(3))
    _lib.check_host(L.pargo_points_read(shp_path.encode(), int(field_index), ctypes.byref(n), _lib._ptr(xs),
                                        _lib._ptr(ys), _lib._ptr(vs)), "pargo_points_read")
    return xs, ys, vs
