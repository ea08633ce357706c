"""The drop-in command lines (slope, flowdird8, idw) on the reference's own
input files, checked against the FP64 oracle.

Bars: slope |gpu - oracle| <= 1e-5 |oracle| + 1e-12 (FP64 arithmetic, float32
storage); D8 codes bit-exact; IDW <= 1e-5 |oracle| + 1e-9; NoData cells
identical.  Each command line is a separate process that opens the GPU."""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import pyoracle as po
from pargo_amd import io

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "pargo_amd", "_build", "bin")
DATA = os.path.join(ROOT, "tests", "golden", "data")
DEM = os.path.join(DATA, "dem.tif")
MOORE = os.path.join(DATA, "moore.nbr")
SAMPLES = os.path.join(DATA, "samples.shp")
MASK = os.path.join(DATA, "mask_500.tif")


def run(args):
    print("run:", " ".join(args), flush=True)
    p = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=100)
    sys.stdout.write(p.stdout)
    return p


def run_ok(args):
    p = run(args)
    assert p.returncode == 0, p.stdout
    return p.stdout


def dem_f32():
    a, info = io.read_geotiff(DEM)
    return a.astype(np.float32), info


@pytest.mark.parametrize("method", ["FD", "FFD", "MD", "SD", "SFD", "TFD", "TFDW"])
def test_slope_cli(tmp_path, method):
    out = str(tmp_path / "slope.tif")
    run_ok([os.path.join(BIN, "slope"), "-elev", DEM, "-nbr", MOORE, "-slp", out, "-mtd", method])
    got, info = io.read_geotiff(out)
    z, dinfo = dem_f32()
    ref = po.slope(z, dinfo["cellsize"], nodata=-9999.0, method=method)
    assert info["transform"] == dinfo["transform"] and info["nodata"] == -9999.0 and info["bits"] == 64
    assert np.array_equal(got == -9999.0, ref == -9999.0)
    v = ref != -9999.0
    assert np.all(np.abs(got[v] - ref[v]) <= 1e-5 * np.abs(ref[v]) + 1e-12)


def test_slope_cli_simple_usage_matches_full(tmp_path):
    a, b = str(tmp_path / "a.tif"), str(tmp_path / "b.tif")
    run_ok([os.path.join(BIN, "slope"), DEM, MOORE, a, "TFD"])
    run_ok([os.path.join(BIN, "slope"), "-elev", DEM, "-nbr", MOORE, "-slp", b, "-mtd", "tfd"])
    assert np.array_equal(io.read_geotiff(a)[0], io.read_geotiff(b)[0])


def test_slope_cli_rejects_other_windows(tmp_path):
    p = run([os.path.join(BIN, "slope"), DEM, os.path.join(DATA, "moore20.nbr"), str(tmp_path / "x.tif")])
    assert p.returncode != 0 and "Moore" in p.stdout


@pytest.mark.parametrize("dirtype", [8, 256])
def test_flowdird8_cli(tmp_path, dirtype):
    out = str(tmp_path / "d8.tif")
    run_ok([os.path.join(BIN, "flowdird8"), DEM, MOORE, out, str(dirtype)])
    got, info = io.read_geotiff(out)
    z, dinfo = dem_f32()
    ref = po.d8(z, nodata=-9999.0, dirtype=dirtype)
    assert info["transform"] == dinfo["transform"] and info["nodata"] == -9999.0
    assert np.array_equal(got, ref)
    assert (got == -1).any()  # dem.tif has flat cells / pits


def test_flowdird8_cli_full_usage_and_bad_type(tmp_path):
    out = str(tmp_path / "d8.tif")
    run_ok([os.path.join(BIN, "flowdird8"), "-elev", DEM, "-nbr", MOORE, "-out", out, "-dirtype", "8"])
    z, _ = dem_f32()
    assert np.array_equal(io.read_geotiff(out)[0], po.d8(z, nodata=-9999.0, dirtype=8))
    p = run([os.path.join(BIN, "flowdird8"), DEM, MOORE, str(tmp_path / "bad.tif"), "7"])
    assert "please input right parameter" in p.stdout and not os.path.exists(str(tmp_path / "bad.tif"))


IDW_ARGS = ["-resolution", "500", "-fieldIndex", "3", "-idwExp", "2", "-searchPointNum", "12", "-searchRange", "0",
            "-blockSize", "5000"]


def idw_reference(mask=True, rows_cap=None):
    xs, ys, vs = io.read_points(SAMPLES, 3)
    m, minfo = io.read_geotiff(MASK)
    m = m.astype(np.int32)
    if rows_cap is not None:
        # the reference reads only the mask rows of the IDW grid's band and
        # caps the lookup at the last of them: repeat that row below it
        m = np.concatenate([m[:rows_cap], np.repeat(m[rows_cap - 1:rows_cap], m.shape[0] - rows_cap, 0)], 0)
    return po.idw(xs, ys, vs, 500.0, k=12, power=2, buffer=0.0, block_size=5000.0,
                  mask=m if mask else None, mask_nodata=int(minfo["nodata"]), nthreads=16)


def check_idw(got, ref):
    assert got.shape == ref.shape
    assert np.array_equal(got == -9999.0, ref == -9999.0)
    v = ref != -9999.0
    assert np.all(np.abs(got[v] - ref[v]) <= 1e-5 * np.abs(ref[v]) + 1e-9)


def test_idw_cli_space(tmp_path):
    out = str(tmp_path / "idw.tif")
    run_ok([os.path.join(BIN, "idw"), "-sample", SAMPLES, "-mask", MASK, "-out", out] + IDW_ARGS + ["-dcmp", "space"])
    got, info = io.read_geotiff(out)
    ref = idw_reference()
    check_idw(got, ref)
    xs, ys, _ = io.read_points(SAMPLES, 3)
    rows, cols, minx, maxy = po.idw_dims(xs, ys, 500.0)
    assert info["transform"] == (minx, 500.0, 0.0, maxy, 0.0, -500.0) and info["nodata"] == -9999.0
    assert os.path.exists(out + ".aux.xml")  # the samples' .prj as GDAL side-car SRS


def test_idw_cli_compute_write_and_read_load(tmp_path):
    load = str(tmp_path / "load.tif")
    a, b = str(tmp_path / "a.tif"), str(tmp_path / "b.tif")
    log = run_ok([os.path.join(BIN, "idw"), "-sample", SAMPLES, "-mask", MASK, "-out", a] + IDW_ARGS +
                 ["-dcmp", "compute", "-writeLoad", load])
    assert "subWorkBR 0 698 699" in log
    xs, ys, _ = io.read_points(SAMPLES, 3)
    m, minfo = io.read_geotiff(MASK)
    ref_load = po.idw_load(xs, ys, 500.0, k=12, block_size=5000.0, granularity=10, mask=m.astype(np.int32),
                           mask_nodata=minfo["nodata"], nodata=-9999.0)
    got_load, linfo = io.read_geotiff(load)
    assert np.array_equal(got_load, ref_load) and linfo["transform"][1] == 5000.0
    # compute mode reads the mask band of the grid's 699 rows only
    check_idw(io.read_geotiff(a)[0], idw_reference(rows_cap=699))
    run_ok([os.path.join(BIN, "idw"), "-sample", SAMPLES, "-mask", MASK, "-out", b] + IDW_ARGS +
           ["-dcmp", "compute", "-readLoad", load])
    assert np.array_equal(io.read_geotiff(a)[0], io.read_geotiff(b)[0])


def test_idw_cli_preliminary_experiment(tmp_path):
    """-dcmp space -writeLoad: the output holds each cell's cost (samples
    examined) instead of its value; -dcmp compute -readLoad balances on it
    and gives the normal IDW grid."""
    work_f, out_f = str(tmp_path / "work.tif"), str(tmp_path / "x.tif")
    run_ok([os.path.join(BIN, "idw"), "-sample", SAMPLES, "-out", work_f] + IDW_ARGS +
           ["-dcmp", "space", "-writeLoad", str(tmp_path / "unused.tif")])
    work, _ = io.read_geotiff(work_f)
    xs, ys, vs = io.read_points(SAMPLES, 3)
    _, cand, _ = po.idw(xs, ys, vs, 500.0, k=12, power=2, buffer=0.0, block_size=5000.0, stats=True)
    assert np.array_equal(work.sum(axis=1).astype(np.int64), cand)
    assert np.all(work == np.round(work)) and work.min() > 0
    run_ok([os.path.join(BIN, "idw"), "-sample", SAMPLES, "-out", out_f] + IDW_ARGS +
           ["-dcmp", "compute", "-readLoad", work_f])
    check_idw(io.read_geotiff(out_f)[0], idw_reference(mask=False))


@pytest.mark.parametrize("converge", [False, True])
def test_scad8_cli_chain(tmp_path, converge):
    """flowdird8 (dirType 8) then scad8 on dem.tif, against the oracle's
    D8 layer and its SCA restatement."""
    d8f, out = str(tmp_path / "d8.tif"), str(tmp_path / "sca.tif")
    run_ok([os.path.join(BIN, "flowdird8"), DEM, MOORE, d8f, "8"])
    run_ok([os.path.join(BIN, "scad8"), d8f, MOORE, out] + (["-converge"] if converge else []))
    got, info = io.read_geotiff(out)
    z, dinfo = dem_f32()
    ref, _ = po.sca_d8(po.d8(z, nodata=-9999.0, dirtype=8), -9999.0, converge)
    assert info["transform"] == dinfo["transform"] and info["nodata"] == -9999.0
    assert np.array_equal(got, ref)
    assert got.max() > 100  # dem.tif has real drainage


def test_scad8_cli_wrong_arguments(tmp_path):
    p = run([os.path.join(BIN, "scad8"), DEM, MOORE])
    assert p.returncode == 0 and "please input right parameter" in p.stdout
