"""Pins the oracle's intensity-ratio compute load (oracle_intensity_load:
Transformation::Operator / run, gpro/transformation.h:155-228, over the
ComputeLayer of gpro/computeLayer.h:208-260).

Closed forms: an all-valid layer gives g*g*validLoad per full granule and
the clipped count at the far edge; NoData cells give nodataLoad; granules
start at the work rectangle's first row/column (the first border row is not
counted, the last one is); several layers add; negative loads are refused.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

ND = -9999.0


def test_all_valid_closed_form():
    z = np.ones((12, 17))
    load = po.intensity_load([z], [ND], 1, 1, 10, 15, 4, 0.0, 2.0)
    assert load.shape == (3, 4)
    # granule rows 1..4, 5..8, 9..12 -> the last is clipped to rows 9..11
    # (the raster's last row is the border row 11, counted)
    rows = [4, 4, 3]
    cols = [4, 4, 4, 4]  # cols 1..4, 5..8, 9..12, 13..16 (16 is the border column)
    assert np.array_equal(load, 2.0 * np.outer(rows, cols))


def test_nodata_cells_use_the_nodata_load():
    z = np.ones((6, 6))
    z[1, 1] = ND
    z[2, 3] = ND + 5e-8  # within 1e-7: NoData
    z[3, 3] = ND + 2e-7  # valid
    load = po.intensity_load([z], [ND], 1, 1, 4, 4, 2, 0.5, 3.0)
    expect = np.zeros((2, 2))
    for r in range(1, 6):
        for c in range(1, 6):
            i, j = (r - 1) // 2, (c - 1) // 2
            if i < 2 and j < 2:
                nd = abs(z[r, c] - ND) <= 1e-7
                expect[i, j] += 0.5 if nd else 3.0
    assert np.array_equal(load, expect)


def test_layers_add_and_first_border_is_skipped():
    a = np.arange(30.0).reshape(5, 6)
    b = np.full((5, 6), ND)
    one = po.intensity_load([a], [ND], 1, 1, 3, 4, 1, 1.0, 1.0)
    two = po.intensity_load([a, b], [ND, ND], 1, 1, 3, 4, 1, 1.0, 1.0)
    assert np.array_equal(two, 2 * one)
    g = po.intensity_load([a], [ND], 1, 1, 3, 4, 5, 0.0, 1.0)
    assert g.shape == (1, 1) and g[0, 0] == 4 * 5  # rows 1..4, cols 1..5


def test_negative_load_is_refused():
    with pytest.raises(ValueError):
        po.intensity_load([np.ones((4, 4))], [ND], 1, 1, 2, 2, 1, -1.0, 1.0)
