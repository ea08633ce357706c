"""Pins of the oracle's IDW compute-load estimate and load-balanced row bands.

The expected values below are traced by hand from the reference's loops
(apps/spatial/idw/idwTransformation.cpp:3-62 and
gpro/deComposition.h:305-393); each case states the trace."""
import numpy as np
import pytest

from oracle import pyoracle as po

# 3 x 3 sample blocks of 10 m on a 30 m x 30 m extent (cell size 1 m, so the
# two anchor samples at (0.5, 0.5) and (29.5, 29.5) make minX = minY = 0,
# maxX = maxY = 30).  Block (r, c) holds COUNTS[r][c] samples; with
# granularity 10 the load grid is 3 x 3 and load cell (r, c) maps to block
# (r, c).
COUNTS = [[2, 0, 1], [4, 5, 0], [1, 3, 2]]


def _samples(counts):
    xs, ys = [], []
    for r in range(3):
        for c in range(3):
            for j in range(counts[r][c]):
                if (r, c) == (2, 0) and j == 0:
                    xs.append(0.5), ys.append(0.5)
                elif (r, c) == (0, 2) and j == 0:
                    xs.append(29.5), ys.append(29.5)
                else:
                    xs.append(10 * c + 4.0 + 0.1 * j), ys.append(25.0 - 10 * r + 0.1 * j)
    return np.array(xs), np.array(ys)


def test_load_grid_hand_traced():
    xs, ys = _samples(COUNTS)
    load = po.idw_load(xs, ys, 1.0, k=4, block_size=10.0, granularity=10)
    # (0,0): 2 < 4; ring 0 adds the centre again -> 4 samples, 1 block: 1 + 2*4
    # (0,1): 0; ring 0 -> 1 block; ring 1 visits only its bottom row
    #        (1,0)=4 -> 2 blocks, 4 samples: 2 + 8
    # (0,2): 1; ring 0 -> 2, 1 block; ring 1 bottom row from (1,1)=5 -> 7: 2 + 14
    # (1,0): 4 >= 4, no ring: 0 + 8;  (1,1): 0 + 10
    # (1,2): 0; ring 0 -> 1 block; ring 1 top row (0,1)=0, (0,2)=1, then the
    #        bottom row (2,1)=3 -> 4 samples after 4 blocks: 4 + 8
    # (2,0): 1; ring 0 -> 2; ring 1 top row (1,0)=4 -> 6 after 2 blocks: 2 + 12
    # (2,1): 3; ring 0 -> 6, 1 block: 1 + 12;  (2,2): 2; ring 0 -> 4: 1 + 8
    assert load.tolist() == [[9, 10, 16], [8, 10, 12], [14, 13, 9]]


def test_load_single_block_closed_form():
    # every sample in one block: K <= n -> 2n; n < K <= 2n -> 1 + 4n
    xs, ys = _samples([[0, 0, 1], [0, 5, 0], [1, 0, 0]])
    n = 7
    load = po.idw_load(xs, ys, 1.0, k=5, block_size=100.0, granularity=10)
    assert np.all(load == 2 * n)
    load = po.idw_load(xs, ys, 1.0, k=12, block_size=100.0, granularity=10)
    assert np.all(load == 1 + 4 * n)


def test_load_never_ends_is_reported():
    # K > 2n on a single block: ring 1 and beyond hold no block
    xs, ys = _samples([[0, 0, 1], [0, 5, 0], [1, 0, 0]])
    with pytest.raises(ValueError, match="-3"):
        po.idw_load(xs, ys, 1.0, k=15, block_size=100.0, granularity=10)


def test_load_mask_gives_nodata():
    xs, ys = _samples(COUNTS)
    mask = np.ones((3, 3), dtype=np.int32)
    mask[1, 1] = 0      # 0 is always masked
    mask[0, 0] = 7      # equal to the mask's NoData
    load = po.idw_load(xs, ys, 1.0, k=4, block_size=10.0, granularity=10, mask=mask, mask_nodata=7.0,
                       nodata=-9999.0)
    assert load[1, 1] == -9999.0 and load[0, 0] == -9999.0
    assert load[2, 2] == 9


def test_load_grid_rounds_up():
    xs, ys = _samples(COUNTS)  # 30 x 30 cells
    assert po.idw_load(xs, ys, 1.0, k=4, block_size=10.0, granularity=7).shape == (5, 5)


def test_bands_uniform_hand_traced():
    # 8 rows of load 3, average 12 per band of 2.  Rows 0-3 reach exactly
    # 12, which leaves the running "closest difference" at 0, read as unset,
    # so row 4 closes the first band: [0, 5).
    assert po.val_row_dcmp(np.full((8, 3), 1.0), -9999.0, 2).tolist() == [0, 5, 8]
    # 4 bands (average 6): row 2 closes [0,3) the same way; then average
    # 15/3 = 5: rows 3-4 give 6, 1 over against 2 under, so [3,5); then
    # average 4.5: rows 5-6 give 6, 1.5 over against 1.5 under -> [5,7).
    assert po.val_row_dcmp(np.full((8, 3), 1.0), -9999.0, 4).tolist() == [0, 3, 5, 7, 8]


def test_bands_heavy_row_hand_traced():
    load = np.array([1, 1, 1, 1, 10, 1, 1, 1.0])[:, None]
    # total 17, average 8.5: rows 0-3 sum 4 (4.5 under), row 4 makes 14
    # (5.5 over) -> the cut goes before row 4: [0,4)
    assert po.val_row_dcmp(load, -9999.0, 2).tolist() == [0, 4, 8]
    # 3 bands, average 17/3: row 4 overshoots (8.33 over vs 1.67 under) ->
    # [0,4); then average 13/2 = 6.5: row 4 alone is 10 (3.5 over, nothing
    # under yet) -> [4,5); the rest is [5,8)
    assert po.val_row_dcmp(load, -9999.0, 3).tolist() == [0, 4, 5, 8]


def test_bands_skip_nodata_and_minus_9999():
    load = np.array([[5, -9999], [5, 1e9], [5, 0], [5, 0]], dtype=np.float64)
    # -9999 and the NoData 1e9 are not load: rows weigh 5 each
    assert po.val_row_dcmp(load, 1e9, 2).tolist() == po.val_row_dcmp(np.full((4, 1), 5.0), 1e9, 2).tolist()


@pytest.mark.parametrize("seed", range(5))
def test_bands_properties(seed):
    rng = np.random.default_rng(seed)
    rows = int(rng.integers(5, 60))
    load = rng.exponential(10.0, size=(rows, 7))
    for nsub in (1, 2, 3, rows):
        f = po.val_row_dcmp(load, -9999.0, nsub)
        assert f[0] == 0 and f[-1] == rows and len(f) == nsub + 1
        assert np.all(np.diff(f) >= 1)  # contiguous, non-empty bands
