"""Pins for the oracle's IDW restatement (oracle/pargo_oracle.c, after
apps/spatial/idw/idwOperator.cpp:5-20,126-169,332-491, 1-rank semantics).

- buffer 0: the ring search returns the exact K nearest samples, so a
  brute-force KNN over all samples pins it (inputs without distance ties);
- buffer > 0: the search stops before ring r once (r-1)*B + m >= buffer
  (m = distance from the cell centre to its block's edges), so the answer is
  the KNN among samples in blocks of Chebyshev ring < r_stop, or 0 if none;
- exact distance ties: hand-built cases of the strict-< replacement rule;
- the extent rule and the mask lookup by hand.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle


def grid_geometry(xs, ys, cs):
    rows, cols, minx, maxy = pyoracle.idw_dims(xs, ys, cs)
    csd = float(np.float32(cs))
    cx = (np.arange(cols) + 0.5) * csd + minx
    cy = maxy - (np.arange(rows) + 0.5) * csd
    return rows, cols, minx, maxy, cx, cy


def brute(xs, ys, vs, cs, k, p, buffer=0.0, block=5000.0):
    rows, cols, minx, maxy, cx, cy = grid_geometry(xs, ys, cs)
    sbr = ((maxy - ys) / block).astype(int)
    sbc = ((xs - minx) / block).astype(int)
    out = np.zeros((rows, cols))
    for r in range(rows):
        for c in range(cols):
            x, y = cx[c], cy[r]
            sel = np.ones(len(xs), bool)
            if buffer > 0:
                br, bc = int((maxy - y) / block), int((x - minx) / block)
                x0, y0 = bc * block + minx, maxy - br * block
                m = min(abs(x - x0), abs(x - (x0 + block)), abs(y - y0), abs(y - (y0 - block)))
                r_stop = 0
                while (r_stop - 1) * block + m < buffer:
                    r_stop += 1
                ring = np.maximum(np.abs(sbr - br), np.abs(sbc - bc))
                sel = ring < r_stop
            if not sel.any():
                out[r, c] = 0.0
                continue
            d = np.hypot(xs[sel] - x, ys[sel] - y)
            d[d == 0] = 1e-7
            o = np.argsort(d, kind="stable")[:k]
            w = 1.0 / d[o] ** p
            out[r, c] = np.sum(vs[sel][o] * w) / np.sum(w)
    return out


def random_samples(n, seed, extent=3000.0):
    rng = np.random.default_rng(seed)
    xs = rng.uniform(0, extent, n)
    ys = rng.uniform(0, extent, n)
    vs = rng.uniform(100, 200, n)
    return xs, ys, vs


@pytest.mark.parametrize("k,p", [(1, 2), (5, 1), (12, 2), (12, 3)])
def test_buffer_zero_is_exact_knn(k, p):
    xs, ys, vs = random_samples(150, 1)
    got = pyoracle.idw(xs, ys, vs, 100.0, k=k, power=p, buffer=0.0, block_size=500.0)
    ref = brute(xs, ys, vs, 100.0, k, p, block=500.0)
    assert np.allclose(got, ref, rtol=1e-12, atol=0)


@pytest.mark.parametrize("buffer", [120.0, 400.0, 900.0])
def test_buffer_ring_rule(buffer):
    xs, ys, vs = random_samples(60, 2)
    got = pyoracle.idw(xs, ys, vs, 100.0, k=12, power=2, buffer=buffer, block_size=500.0)
    ref = brute(xs, ys, vs, 100.0, 12, 2, buffer=buffer, block=500.0)
    assert np.allclose(got, ref, rtol=1e-12, atol=0)
    if buffer == 120.0:
        assert (got == 0).any()  # cells whose first ring has no samples


def test_single_sample_constant():
    xs, ys, vs = np.array([1000.0]), np.array([2000.0]), np.array([42.5])
    out = pyoracle.idw(np.r_[xs, 1500.0], np.r_[ys, 2600.0], np.r_[vs, 42.5], 50.0, k=1, power=2)
    # (v * w) / w rounds twice, so "constant" holds to an ulp
    assert np.allclose(out, 42.5, rtol=2e-16, atol=0)


def test_coincident_sample_uses_eps_distance():
    # samples on cell centres: weight 1/1e-14 swamps the others
    xs = np.array([15.0, 75.0, 45.0])
    ys = np.array([15.0, 75.0, 45.0])
    vs = np.array([1.0, 3.0, 7.0])
    out = pyoracle.idw(xs, ys, vs, 30.0, k=3, power=2)
    rows, cols, minx, maxy, cx, cy = grid_geometry(xs, ys, 30.0)
    for x, y, v in zip(xs, ys, vs):
        r = int(np.where(cy == y)[0][0])
        c = int(np.where(cx == x)[0][0])
        assert out[r, c] == pytest.approx(v, rel=1e-10)


def test_extent_rule():
    # minX = 100 - 15 = 85; (200.5-85)/30 = 3.85 -> 3, not exact -> 4 cols;
    # minY = 50 - 15 = 35; (110-35)/30 = 2.5 -> 2 -> 3 rows; maxY = 35 + 90.
    rows, cols, minx, maxy = pyoracle.idw_dims([100.0, 200.5], [50.0, 110.0], 30.0)
    assert (rows, cols, minx, maxy) == (3, 4, 85.0, 125.0)
    # anchors half a cell inside a 4096-cell square give exactly 4096 x 4096
    assert pyoracle.idw_dims([15.0, 122865.0], [15.0, 122865.0], 30.0) == (4096, 4096, 0.0, 122880.0)


def tie_layout(extra):
    """cs 1, block 100, a 1000x1000 grid fixed by two far anchors."""
    xs = [0.5, 999.5] + [e[0] for e in extra]
    ys = [0.5, 999.5] + [e[1] for e in extra]
    vs = [-1.0, -1.0] + [e[2] for e in extra]
    return np.array(xs), np.array(ys), np.array(vs)


def test_tie_keeps_first_visited():
    # K=1, two samples 5 m from the centre (550.5, 450.5) of cell (549, 550)
    a, b = (555.5, 450.5, 10.0), (545.5, 450.5, 20.0)
    out = pyoracle.idw(*tie_layout([a, b]), 1.0, k=1, power=2, block_size=100.0)
    assert out[549, 550] == 10.0
    out = pyoracle.idw(*tie_layout([b, a]), 1.0, k=1, power=2, block_size=100.0)
    assert out[549, 550] == 20.0


def test_replacement_evicts_first_max_slot():
    # A (d=5), B (d=5), C (d=3) in one block, K=2: slots [A, B], C replaces
    # the first slot holding the maximum -> {C, B}.
    a, b, c = (555.5, 450.5, 10.0), (545.5, 450.5, 20.0), (550.5, 453.5, 40.0)
    out = pyoracle.idw(*tie_layout([a, b, c]), 1.0, k=2, power=2, block_size=100.0)
    expect = (40.0 / 9 + 20.0 / 25) / (1 / 9 + 1 / 25)
    assert out[549, 550] == pytest.approx(expect, rel=1e-14)


def test_no_sample_inside_buffer_gives_zero():
    a = (555.5, 450.5, 10.0)
    out = pyoracle.idw(*tie_layout([a]), 1.0, k=1, power=2, buffer=10.0, block_size=100.0)
    assert out[800, 250] == 0.0  # block (8, 2): empty, and ring 1 starts beyond 10 m
    assert out[549, 550] == 10.0


def test_fewer_samples_than_k_without_buffer_is_an_error():
    xs, ys, vs = random_samples(5, 3)
    with pytest.raises(ValueError):
        pyoracle.idw(xs, ys, vs, 100.0, k=12, power=2, buffer=0.0)
    out = pyoracle.idw(xs, ys, vs, 100.0, k=12, power=2, buffer=1e6, block_size=500.0)
    ref = brute(xs, ys, vs, 100.0, 12, 2, buffer=1e6, block=500.0)
    assert np.allclose(out, ref, rtol=1e-12)


def test_mask_lookup():
    xs, ys, vs = random_samples(40, 4, extent=900.0)
    rows, cols, _, _ = pyoracle.idw_dims(xs, ys, 100.0)
    mr, mc = 5, 7
    rng = np.random.default_rng(5)
    mask = rng.integers(0, 3, size=(mr, mc)).astype(np.int32)  # 0 masked, 2 = mask NoData
    out = pyoracle.idw(xs, ys, vs, 100.0, k=4, power=2, block_size=300.0, mask=mask, mask_nodata=2)
    ref = brute(xs, ys, vs, 100.0, 4, 2, block=300.0)
    for r in range(rows):
        for c in range(cols):
            m = mask[min(int(mr / rows * r), mr - 1), min(int(mc / cols * c), mc - 1)]
            if m in (0, 2):
                assert out[r, c] == -9999.0
            else:
                assert out[r, c] == pytest.approx(ref[r, c], rel=1e-12)


def test_candidate_counts_cover_rings():
    xs, ys, vs = random_samples(200, 6)
    out, cand, blk = pyoracle.idw(xs, ys, vs, 100.0, k=12, power=2, block_size=500.0, stats=True)
    rows, cols = out.shape
    assert cand.sum() >= 12 * rows * cols
    assert (blk > 0).all()
