"""GeoTIFF and shapefile codecs of libpargo (host only, no GPU).

Expected values come from tests/golden/io_expected.json, written by
tests/golden/make_fixtures.py with Pillow's TIFF decoder and an independent
struct/DBF parse of the reference's own input files."""
import hashlib
import json
import os

import numpy as np
import pytest

from pargo_amd import io

GOLD = os.path.join(os.path.dirname(__file__), "golden")
DATA = os.path.join(GOLD, "data")
EXP = json.load(open(os.path.join(GOLD, "io_expected.json")))


@pytest.mark.parametrize("name", ["dem.tif", "mask_500.tif"])
def test_geotiff_matches_independent_decoder(name):
    a, info = io.read_geotiff(os.path.join(DATA, name))
    e = EXP[name]
    assert (info["rows"], info["cols"]) == (e["rows"], e["cols"])
    assert list(info["transform"]) == e["transform"]
    assert info["nodata"] == e["nodata"]
    assert hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest() == e["sha256_f64"]


def test_dem_header_as_the_reference_reads_it():
    # dem.tif: strip float32, GDAL_NODATA "-9999"; MetaData::cellSize is
    # float(pTransform[1]) (gpro/rasterLayer.h:770)
    info = io.geotiff_info(os.path.join(DATA, "dem.tif"))
    assert info["bits"] == 32 and info["sample_format"] == 3
    assert info["cellsize"] == 90.0 and info["nodata"] == -9999.0


def test_mask_is_lzw_tiled_uint8():
    a, info = io.read_geotiff(os.path.join(DATA, "mask_500.tif"))
    assert info["bits"] == 8 and info["sample_format"] == 1
    assert set(np.unique(a)) == {0.0, 1.0}


@pytest.mark.parametrize("dtype,npdt", [("float64", np.float64), ("float32", np.float32), ("int32", np.int32),
                                        ("uint8", np.uint8)])
def test_geotiff_write_read_round_trip(tmp_path, dtype, npdt):
    rng = np.random.default_rng(7)
    a = rng.uniform(0, 250, size=(37, 53))
    if dtype in ("int32", "uint8"):
        a = np.floor(a)
    a = a.astype(npdt).astype(np.float64)
    gt = (437168.40625, 30.0, 0.0, 5428606.5, 0.0, -30.0)
    p = str(tmp_path / ("o_%s.tif" % dtype))
    io.write_geotiff(p, a, gt, nodata=-9999.0, dtype=dtype, georef_from=os.path.join(DATA, "dem.tif"))
    b, info = io.read_geotiff(p)
    assert np.array_equal(a, b)
    assert info["transform"] == gt and info["nodata"] == -9999.0
    assert info["bits"] == {"float64": 64, "float32": 32, "int32": 32, "uint8": 8}[dtype]


def test_geotiff_write_keeps_nodata_text_exact(tmp_path):
    p = str(tmp_path / "nd.tif")
    nd = -3.4028234663852886e38  # float32 lowest, a common NoData
    io.write_geotiff(p, np.zeros((3, 4)), (0, 1, 0, 0, 0, -1), nodata=nd)
    assert io.geotiff_info(p)["nodata"] == nd


def test_geotiff_missing_file_raises():
    from pargo_amd._lib import PargoError
    with pytest.raises(PargoError):
        io.read_geotiff(os.path.join(DATA, "no_such.tif"))


def test_points_match_independent_parse():
    xs, ys, vs = io.read_points(os.path.join(DATA, "samples.shp"), 3)
    e = EXP["samples.shp"]
    assert len(xs) == e["n"]
    arr = np.array([xs, ys, vs], dtype=np.float64)
    assert hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest() == e["sha256_xyv"]
    assert [xs.min(), ys.min(), xs.max(), ys.max()] == e["bbox"]


def test_points_bad_field_index_raises():
    from pargo_amd._lib import PargoError
    with pytest.raises(PargoError):
        io.read_points(os.path.join(DATA, "samples.shp"), 17)
