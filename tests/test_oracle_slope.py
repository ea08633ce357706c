"""Pins for the oracle's slope restatement (oracle/pargo_oracle.c, after
apps/morphology/slope/slopeOperator.cpp:51-139).

Each check is independent of the oracle's code: closed forms on planes, a
library Sobel operator (Horn's weights), and the two documented quirks of
the reference derived by hand (TFD subtracts d[6] twice in dy; MD divides by
the last neighbour's distance cellSize*sqrt(2)).
"""
import math

import numpy as np
import pytest
from scipy import ndimage

from oracle import pyoracle

ND = -9999.0
CS = 30.0


def plane(rows, cols, gx, gy, z0=1000.0, cs=CS):
    """z rises by gx per metre to the east and gy per metre to the north
    (row 0 is the northern edge)."""
    r = np.arange(rows)[:, None]
    c = np.arange(cols)[None, :]
    return (z0 + gx * cs * c - gy * cs * r).astype(np.float64)


GOOD_ON_PLANES = ["FD", "FFD", "SD", "SFD", "TFDW"]


@pytest.mark.parametrize("method", GOOD_ON_PLANES)
@pytest.mark.parametrize("gx,gy", [(0.25, -0.5), (1.0, 0.0), (0.0, 0.75), (-0.125, 0.375)])
def test_plane_gives_gradient_magnitude(method, gx, gy):
    z = plane(9, 11, gx, gy)
    s = pyoracle.slope(z, CS, ND, method)
    inner = s[1:-1, 1:-1]
    assert np.allclose(inner, math.hypot(gx, gy), rtol=1e-14, atol=0)


def test_tfd_quirk_on_east_west_plane():
    # dy = (d2 + d1 + d0 - d6 - d7 - d6) / 6cs: on z = a*x the d[6]-twice
    # term leaves dy = 2a*cs / 6cs = a/3, so slope = a*sqrt(10)/3.
    a = 0.5
    s = pyoracle.slope(plane(7, 7, a, 0.0), CS, ND, "TFD")
    assert np.allclose(s[1:-1, 1:-1], a * math.sqrt(10) / 3, rtol=1e-14)
    # on z = b*y the quirk is invisible: slope = |b|
    s = pyoracle.slope(plane(7, 7, 0.0, 0.4), CS, ND, "TFD")
    assert np.allclose(s[1:-1, 1:-1], 0.4, rtol=1e-14)


def test_md_quirk_divides_by_diagonal_distance():
    # Steepest drop on z = x (gx = 1) is cs towards W, NW and SW; the
    # reference divides by cs*sqrt(2) (the SE neighbour's distance).
    s = pyoracle.slope(plane(6, 6, 1.0, 0.0), CS, ND, "MD")
    assert np.allclose(s[1:-1, 1:-1], 1 / math.sqrt(2), rtol=1e-14)
    # gx = gy = 1: the SW neighbour drops 2cs -> 2/sqrt(2)
    s = pyoracle.slope(plane(6, 6, 1.0, 1.0), CS, ND, "MD")
    assert np.allclose(s[1:-1, 1:-1], math.sqrt(2), rtol=1e-14)
    # a pit or a plateau has no drop: 0
    z = np.full((5, 5), 7.0)
    z[2, 2] = 3.0
    assert pyoracle.slope(z, CS, ND, "MD")[2, 2] == 0.0


@pytest.mark.parametrize("method", list(pyoracle.METHODS))
def test_constant_surface_is_flat(method):
    s = pyoracle.slope(np.full((6, 9), 123.5), CS, ND, method)
    if method == "TFDW":
        # left-to-right FP64 sums with sqrt(2.0)*d terms do not cancel
        # exactly: the reference reports ~1e-17 on a flat surface
        assert np.all(np.abs(s[1:-1, 1:-1]) < 1e-14)
    else:
        assert np.all(s[1:-1, 1:-1] == 0.0)


@pytest.mark.parametrize("method", list(pyoracle.METHODS))
def test_border_and_nodata(method):
    rng = np.random.default_rng(1)
    z = rng.uniform(0, 100, size=(12, 14))
    z[5, 6] = ND
    s = pyoracle.slope(z, CS, ND, method)
    assert np.all(s[0, :] == ND) and np.all(s[-1, :] == ND)
    assert np.all(s[:, 0] == ND) and np.all(s[:, -1] == ND)
    assert np.all(s[4:7, 5:8] == ND)  # every window holding (5, 6)
    others = np.ones_like(s, dtype=bool)
    others[[0, -1], :] = False
    others[:, [0, -1]] = False
    others[4:7, 5:8] = False
    assert np.all(s[others] != ND)


def test_nodata_epsilon_is_1e_7():
    z = np.full((5, 5), 50.0)
    z[2, 2] = ND + 0.5e-7  # within 1e-7 of noData: NoData
    assert pyoracle.slope(z, CS, ND)[2, 2] == ND
    z[2, 2] = ND + 2e-7  # outside: an ordinary (very low) value
    assert pyoracle.slope(z, CS, ND)[2, 2] != ND


def test_horn_matches_library_sobel():
    """Horn's dz/dx and dz/dy are the Sobel responses / (8 cs)."""
    rng = np.random.default_rng(2)
    z = rng.uniform(0, 2000, size=(40, 57)).astype(np.float32).astype(np.float64)
    s = pyoracle.slope(z, CS, ND, "FD")
    sx = ndimage.sobel(z, axis=1, mode="nearest")
    sy = ndimage.sobel(z, axis=0, mode="nearest")
    ref = np.hypot(sx, sy) / (8 * CS)
    assert np.allclose(s[1:-1, 1:-1], ref[1:-1, 1:-1], rtol=1e-12, atol=1e-12)


def test_cellsize_is_float_rounded():
    # MetaData::cellSize is float: 0.1 is used as float32(0.1).
    z = plane(5, 5, 1.0, 0.0, cs=float(np.float32(0.1)))
    s = pyoracle.slope(z, 0.1, ND, "SD")
    assert np.allclose(s[1:-1, 1:-1], 1.0, rtol=1e-15)


def test_row_bands_match_single_thread():
    """One band per thread (the MPI_OpenMP_Type sweep) equals the serial
    sweep, cell for cell."""
    rng = np.random.default_rng(3)
    z = rng.uniform(0, 2000, size=(101, 67))
    a = pyoracle.slope(z, CS, ND, "FD", nthreads=1)
    b = pyoracle.slope(z, CS, ND, "FD", nthreads=7)
    assert np.array_equal(a, b)


def test_smpl1d_dcmp_bands():
    # deComposition.h:83-115: 10 rows over 3 parts -> 4, 3, 3
    assert [pyoracle.smpl1d_dcmp(1, 10, 3, i) for i in range(3)] == [(1, 4), (5, 7), (8, 10)]
    assert [pyoracle.smpl1d_dcmp(0, 7, 4, i) for i in range(4)] == [(0, 1), (2, 3), (4, 5), (6, 7)]
