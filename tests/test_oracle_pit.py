"""Pins the oracle's pitremove restatement (oracle/pargo_oracle.c,
oracle_pit_remove: PitRemoveOperator, apps/hydrology/pitremove/
pitRemoveOperator.cpp:26-72, swept until a sweep assigns nothing).

Checks:
  * hand-derived values: an outlet ring keeps the DEM, a flat interior ring
    rises by one gap, a pit by two gaps (all with the reference's double
    additions);
  * an independent solver: the reference ends at the solution of
        W(c) = dem(c)                            if dem(c) >= W(n) + gap or W(n) is NoData for a neighbour n
        W(c) = min(10000, min_n (W(n) + gap))    otherwise
    (W(c) = 10000 where dem(c) >= 10000).  The increments are positive, so
    the solution is unique and a priority flood (cells settled in order of
    W, Dijkstra-style) computes it without any sweep order;
  * cells never visited (border) and NoData cells hold NoData.
"""
import heapq

import numpy as np
import pytest

from oracle import pyoracle as po

ND = -9999.0
GAP = 0.0005


def priority_flood(dem, nd=ND):
    rows, cols = dem.shape
    W = np.full(dem.shape, nd)
    valid = np.zeros(dem.shape, bool)
    for r in range(1, rows - 1):
        for c in range(1, cols - 1):
            if abs(dem[r, c] - nd) > 1e-6:
                valid[r, c] = True
                W[r, c] = 10000.0
    done = np.zeros(dem.shape, bool)
    heap = []
    for r in range(1, rows - 1):
        for c in range(1, cols - 1):
            if not valid[r, c] or dem[r, c] >= 10000.0:
                continue
            # next to NoData or the border: W = dem
            if any(not valid[r + dr, c + dc] for dr in (-1, 0, 1) for dc in (-1, 0, 1) if (dr, dc) != (0, 0)):
                W[r, c] = dem[r, c]
                heapq.heappush(heap, (W[r, c], r, c))
    # frozen cells (dem >= 10000) keep 10000 and take part as neighbours
    for r in range(1, rows - 1):
        for c in range(1, cols - 1):
            if valid[r, c] and dem[r, c] >= 10000.0:
                heapq.heappush(heap, (10000.0, r, c))
    while heap:
        w, r, c = heapq.heappop(heap)
        if done[r, c] or w != W[r, c]:
            continue
        done[r, c] = True
        s = w + GAP
        for dr in (-1, 0, 1):
            for dc in (-1, 0, 1):
                rr, cc = r + dr, c + dc
                if (dr, dc) == (0, 0) or not valid[rr, cc] or done[rr, cc] or dem[rr, cc] >= 10000.0:
                    continue
                cand = dem[rr, cc] if dem[rr, cc] >= s else min(10000.0, s)
                if cand < W[rr, cc]:
                    W[rr, cc] = cand
                    heapq.heappush(heap, (cand, rr, cc))
    return W


def test_rings_closed_form():
    dem = np.full((7, 7), ND)
    dem[1:-1, 1:-1] = 10.0
    dem[3, 3] = 1.0
    W, sweeps, ok = po.pit_remove(dem)
    assert ok
    assert np.all(W[1:-1, 1:-1][[0, -1], :] == 10.0) and np.all(W[1:-1, [1, -2]] == 10.0)
    inner = [(2, 2), (2, 3), (2, 4), (3, 2), (3, 4), (4, 2), (4, 3), (4, 4)]
    assert all(W[p] == 10.0 + GAP for p in inner)
    assert W[3, 3] == (10.0 + GAP) + GAP
    assert np.all(W[0] == ND) and np.all(W[:, -1] == ND)


def test_outlets_keep_the_dem():
    dem = np.full((5, 6), ND)
    dem[1:-1, 1:-1] = np.arange(12, dtype=float).reshape(3, 4)
    W, _, _ = po.pit_remove(dem)
    assert np.array_equal(W[1:-1, 1:-1], dem[1:-1, 1:-1])  # every interior cell touches the border


@pytest.mark.parametrize("seed", range(10))
def test_matches_priority_flood(seed):
    rng = np.random.default_rng(seed)
    rows, cols = int(rng.integers(3, 26)), int(rng.integers(3, 26))
    dem = np.round(rng.random((rows, cols)) * 20, 1).astype(np.float32).astype(np.float64)
    if seed % 2:
        dem[rng.random((rows, cols)) < 0.08] = ND  # NoData cells are outlets for their neighbours
    if seed % 3 == 0:
        dem[rng.integers(0, rows), rng.integers(0, cols)] = 12000.0  # above the start value: frozen at 10000
    W, sweeps, ok = po.pit_remove(dem)
    assert ok
    assert np.array_equal(W, priority_flood(dem)), (rows, cols)


def test_plateau_is_tilted_toward_its_outlet():
    dem = np.full((4, 10), ND)
    dem[1:-1, 1:-1] = 5.0
    dem[1:-1, 1] = 5.0
    W, _, _ = po.pit_remove(dem)
    assert np.array_equal(W, priority_flood(dem))


def test_sweep_count_grows_with_upstream_propagation():
    # a walled plateau whose only low outlet is the bottom-right corner: its
    # tilt has to travel up-left, against the row-major sweep, about one
    # cell per sweep
    dem = np.full((40, 40), ND)
    dem[1:-1, 1:-1] = 100.0
    dem[2:-2, 2:-2] = 50.0
    dem[38, 38] = 0.0
    W, sweeps, ok = po.pit_remove(dem)
    assert ok and sweeps > 30
    assert W[2, 2] > W[37, 37] == 50.0
    assert np.array_equal(W, priority_flood(dem))


def test_sweep_limit_reports_not_converged():
    dem = np.full((30, 30), ND)
    dem[1:-1, 1:-1] = 100.0
    dem[2:-2, 2:-2] = 50.0
    _, sweeps, ok = po.pit_remove(dem, max_sweeps=2)
    assert not ok and sweeps == 2
