"""The tile schedule of the device's depression filling reaches the oracle's
result.

pargo_pit_remove runs passes over tiles: a tile sweeps its cells with the
reference's update (pitRemoveOperator.cpp:46-70) against a halo loaded when
it starts, then, for each of its edges on which a cell changed, schedules
the three neighbour tiles on that side for the next pass.  Pass 0 holds the
tiles with NoData or border cells in their one-cell halo; a tile stamped for
the next pass while it waits in the current list is deferred.  This test runs that schedule on the
CPU with small tiles, so that many tiles and corners take part, and
requires the oracle's values: a schedule that leaves out a tile that could
still change stops short of the fixed point.
"""
import numpy as np
import pytest

from oracle import pyoracle

ND = -9999.0
GAP = 0.0005
START = 10000.0


def update(W, D, live, r, c):
    """One update of cell (r, c); True if it changed."""
    if not live[r, c]:
        return False
    w, d = W[r, c], D[r, c]
    if w <= d:
        return False
    to_dem, m = False, w
    for dr in (-1, 0, 1):
        for dc in (-1, 0, 1):
            if dr == 0 and dc == 0:
                continue
            wn = W[r + dr, c + dc]
            s = wn + GAP
            if d >= s or abs(wn - ND) < 1e-6:
                to_dem = True
            elif m > s:
                m = s
    nw = d if to_dem else m
    if nw == w:
        return False
    W[r, c] = nw
    return True


def schedule_fill(z, th, tw):
    rows, cols = z.shape
    D = z.astype(np.float64)
    interior = np.zeros((rows, cols), bool)
    interior[1:-1, 1:-1] = True
    live = interior & (np.abs(D - ND) > 1e-6)
    W = np.where(live, START, ND)
    ty_n, tx_n = (rows + th - 1) // th, (cols + tw - 1) // tw
    stamp = -np.ones(ty_n * tx_n, int)
    lists = [[], []]

    def sched(t, p):
        if stamp[t] != p:
            stamp[t] = p
            lists[p & 1].append(t)

    for r in range(rows):
        for c in range(cols):
            if not live[r, c]:
                for ty in range(max(r - 1, 0) // th, min(r + 1, rows - 1) // th + 1):
                    for tx in range(max(c - 1, 0) // tw, min(c + 1, cols - 1) // tw + 1):
                        sched(ty * tx_n + tx, 0)
    nbrs = [(-1, -1), (-1, 0), (-1, 1), (0, -1), (0, 1), (1, -1), (1, 0), (1, 1)]
    p = 0
    while lists[p & 1]:
        cur, lists[p & 1] = lists[p & 1], []
        for t in cur:
            if stamp[t] != p:
                continue  # deferred to a later pass
            ty, tx = divmod(t, tx_n)
            r0, c0 = ty * th, tx * tw
            r1, c1 = min(r0 + th, rows), min(c0 + tw, cols)
            # the halo as loaded: a frame of the current W around the tile
            halo = {}
            for r in range(r0 - 1, r1 + 1):
                for c in range(c0 - 1, c1 + 1):
                    if not (r0 <= r < r1 and c0 <= c < c1):
                        halo[(r, c)] = W[r, c] if 0 <= r < rows and 0 <= c < cols else ND
            # the tile's own cells and its halo snapshot, in a private copy
            Wl = np.full((r1 - r0 + 2, c1 - c0 + 2), ND)
            for (r, c), v in halo.items():
                Wl[r - r0 + 1, c - c0 + 1] = v
            Wl[1:-1, 1:-1] = W[r0:r1, c0:c1]
            Dl = np.zeros_like(Wl)
            Dl[1:-1, 1:-1] = D[r0:r1, c0:c1]
            Ll = np.zeros(Wl.shape, bool)
            Ll[1:-1, 1:-1] = live[r0:r1, c0:c1]
            touched = np.zeros(Wl.shape, bool)
            changed = True
            while changed:
                changed = False
                for rr in list(range(1, r1 - r0 + 1)) + list(range(r1 - r0, 0, -1)):
                    for cc in range(1, c1 - c0 + 1):
                        if update(Wl, Dl, Ll, rr, cc):
                            changed = touched[rr, cc] = True
            W[r0:r1, c0:c1] = Wl[1:-1, 1:-1]
            # changed cells on the tile's full-size edges (a partial tile at
            # the grid's end has no neighbour past it)
            top = bool(touched[1].any())
            bottom = bool(touched[th].any()) if r1 - r0 == th else False
            left = bool(touched[:, 1].any())
            right = bool(touched[:, tw].any()) if c1 - c0 == tw else False
            for dy, dx in nbrs:
                need = (dy == -1 and top) or (dy == 1 and bottom) or (dx == -1 and left) or (dx == 1 and right)
                nty, ntx = ty + dy, tx + dx
                if need and 0 <= nty < ty_n and 0 <= ntx < tx_n:
                    sched(nty * tx_n + ntx, p + 1)
        p += 1
    return W, p


@pytest.mark.parametrize("seed", range(6))
def test_tile_schedule_reaches_the_oracle(seed):
    rng = np.random.default_rng(700 + seed)
    rows, cols = int(rng.integers(12, 30)), int(rng.integers(12, 40))
    z = np.round(rng.random((rows, cols)) * 20, 2).astype(np.float32)
    z[rng.random((rows, cols)) < 0.04] = ND
    if seed % 2:
        z[rng.random((rows, cols)) < 0.03] = 12000.0  # frozen at the start value
    ref, _, ok = pyoracle.pit_remove(z, ND)
    assert ok
    got, passes = schedule_fill(z, 4, 8)
    assert np.array_equal(got, ref), np.argwhere(got != ref)[:5].tolist()


def test_walled_bowl_with_an_outlet_at_a_tile_corner():
    # a bowl inside a wall, draining to one NoData cell at the corner of four
    # tiles, so that its front leaves that cell into each of them
    z = np.full((26, 34), 50.0, np.float32)
    z[1:-1, 1:-1] = 90.0
    z[3:-3, 3:-3] = 10.0
    z[8, 16] = ND
    ref, _, ok = pyoracle.pit_remove(z, ND)
    assert ok
    got, _ = schedule_fill(z, 4, 8)
    assert np.array_equal(got, ref)


def frozen_grid(rows, cols):
    z = np.full((rows, cols), 12000.0, np.float32)  # never updated
    return z


def test_seed_reaches_the_tile_across_an_edge():
    # everything frozen but a pit in tile (1, 1), whose only outlet is a NoData
    # cell in the last column of tile (1, 0): only pass 0's seeding of the
    # tiles whose halo holds that cell can start the pit's tile
    z = frozen_grid(12, 24)
    z[4:7, 8:11] = 5.0
    z[5, 7] = ND
    ref, _, ok = pyoracle.pit_remove(z, ND)
    assert ok and ref[5, 8] == 5.0
    got, _ = schedule_fill(z, 4, 8)
    assert np.array_equal(got, ref)


def test_a_changed_corner_reaches_the_diagonal_tile():
    # everything frozen but tile (1, 1)'s last cell, drained by a NoData cell
    # inside that tile, and a pit in tile (2, 2) starting at its first cell:
    # the pit changes only once tile (1, 1)'s corner has, through the
    # diagonal neighbour
    z = frozen_grid(14, 28)
    z[7, 15] = 5.0
    z[6, 14] = ND
    z[8:11, 16:19] = 5.0
    ref, _, ok = pyoracle.pit_remove(z, ND)
    assert ok and ref[8, 16] != 10000.0
    got, _ = schedule_fill(z, 4, 8)
    assert np.array_equal(got, ref)
