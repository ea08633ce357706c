"""Pins the oracle's scad8 restatement (oracle/pargo_oracle.c, oracle_sca_d8).

The oracle replays SCAOperator::Operator (apps/hydrology/scad8/
scaOperator.cpp:37-131) sweep by sweep.  These tests check it against:
  * closed forms: a grid draining straight south accumulates row by row in
    a single sweep; a grid draining north gets 1, 2, 1, 1, ... because the
    single sweep only moves area forward in row-major order;
  * brute force: when every code points forward (E, SW, S, SE) one sweep
    already finishes the job, so the area equals the number of cells whose
    flow path passes through the cell; the converging variant equals that
    brute force for ANY acyclic code grid;
  * an independent formulation of one sweep by position-ordered recursion
    (done / available, written from the loop's semantics, not its order);
  * hand-traced windows for the mutual-flow exclusion, border inflows and
    the last-cell rule (a NoData last cell adds its inflows to NoData).
"""
import random

import numpy as np
import pytest

from oracle import pyoracle as po

ND = -9999.0
OFF = [(-1, -1), (-1, 0), (-1, 1), (0, -1), (0, 0), (0, 1), (1, -1), (1, 0), (1, 1)]


def bordered(inner, nd=ND):
    a = np.full((inner.shape[0] + 2, inner.shape[1] + 2), nd)
    a[1:-1, 1:-1] = inner
    return a


def target(d8, r, c):
    v = d8[r, c]
    if v in (0, 1, 2, 3, 5, 6, 7, 8):
        dr, dc = OFF[int(v)]
        return r + dr, c + dc
    return None


def brute_upstream(d8, nd=ND):
    """Cells draining through each interior cell (itself included), by
    following every interior cell's path; needs an acyclic grid without
    mutual pairs and without inflow from border or NoData cells."""
    rows, cols = d8.shape
    out = np.full(d8.shape, nd)
    for r in range(1, rows - 1):
        for c in range(1, cols - 1):
            if abs(d8[r, c] - nd) >= 1e-6:
                out[r, c] = 0
    for r in range(1, rows - 1):
        for c in range(1, cols - 1):
            if abs(d8[r, c] - nd) < 1e-6:
                continue
            cur, seen = (r, c), set()
            while cur is not None and 1 <= cur[0] <= rows - 2 and 1 <= cur[1] <= cols - 2:
                if abs(d8[cur] - nd) < 1e-6:
                    break
                assert cur not in seen
                seen.add(cur)
                out[cur] += 1
                cur = target(d8, *cur)
    return out


def model_one_sweep(d8, nd=ND):
    """One accumulation sweep stated as a recursion on row-major position."""
    rows, cols = d8.shape
    interior = lambda r, c: 1 <= r <= rows - 2 and 1 <= c <= cols - 2
    valid = lambda r, c: abs(d8[r, c] - nd) > 1e-6
    nodata = lambda r, c: abs(d8[r, c] - nd) < 1e-6  # neither: NaN, area 1, no inflow

    def inflows(r, c):
        res = []
        for k, (dr, dc) in enumerate(OFF):
            if k == 4:
                continue
            u = (r + dr, c + dc)
            if d8[u] == 8 - k and d8[r, c] + (8 - k) != 8:
                res.append((k, u))
        return res

    done, val = {}, {}
    for r in range(1, rows - 1):
        for c in range(1, cols - 1):
            if nodata(r, c):
                continue
            if not valid(r, c):
                done[(r, c)], val[(r, c)] = True, 1.0
                continue
            ins = inflows(r, c)
            ok, v = True, 1.0
            for k, u in ins:
                src = interior(*u) and valid(*u) and not inflows(*u)
                if src:
                    v += 1.0
                elif interior(*u) and k < 4 and done[u]:
                    v += val[u]
                else:
                    ok = False
            done[(r, c)], val[(r, c)] = ok, v
    out = np.full(d8.shape, nd)
    for (r, c), v in val.items():
        out[r, c] = v
    # last work cell: a NoData cell there adds every finished inflow
    lr, lc = rows - 2, cols - 2
    if rows >= 3 and cols >= 3 and not valid(lr, lc):
        s = nd if nodata(lr, lc) else 1.0
        for k, (dr, dc) in enumerate(OFF):
            u = (lr + dr, lc + dc)
            if k != 4 and d8[u] == 8 - k and u in done and done[u]:
                s += val[u]
        out[lr, lc] = s
    return out


def test_south_closed_form():
    d8 = bordered(np.full((7, 5), 7.0))
    s, sweeps = po.sca_d8(d8)
    assert sweeps == 3
    for r in range(1, 8):
        assert np.all(s[r, 1:-1] == r)
    assert np.all(s[0] == ND) and np.all(s[-1] == ND) and np.all(s[:, 0] == ND) and np.all(s[:, -1] == ND)


def test_north_single_sweep_quirk():
    d8 = bordered(np.full((6, 4), 1.0))
    s, _ = po.sca_d8(d8)
    expect = np.ones((6, 4))
    expect[-2] = 2  # only the last row's sources reach their northern neighbour
    assert np.array_equal(s[1:-1, 1:-1], expect)
    full, sweeps = po.sca_d8(d8, converge=True)
    assert np.array_equal(full[1:-1, 1:-1], np.repeat(np.arange(6, 0, -1.0)[:, None], 4, axis=1))
    assert sweeps > 3


@pytest.mark.parametrize("seed", range(6))
def test_forward_codes_match_brute_force(seed):
    rng = np.random.default_rng(seed)
    inner = rng.choice([5.0, 6.0, 7.0, 8.0, -1.0], size=(9, 11))
    d8 = bordered(inner)
    # forward codes never come from the (NoData) border; cells draining off
    # the grid simply end their path
    s, _ = po.sca_d8(d8)
    assert np.array_equal(s, brute_upstream(d8))


def acyclic_codes(rng, rows, cols, nodata_frac=0.0):
    """Codes from a random height field (strict descent, so no cycles)."""
    h = rng.permutation(rows * cols).reshape(rows, cols).astype(float)
    d8 = np.full((rows, cols), -1.0)
    for r in range(rows):
        for c in range(cols):
            best, bk = 0.0, -1
            for k, (dr, dc) in enumerate(OFF):
                rr, cc = r + dr, c + dc
                if k == 4 or not (0 <= rr < rows and 0 <= cc < cols):
                    continue
                drop = h[r, c] - h[rr, cc]
                if drop > best:
                    best, bk = drop, k
            d8[r, c] = bk
    if nodata_frac:
        d8[rng.random((rows, cols)) < nodata_frac] = ND
    d8[0, :] = d8[-1, :] = d8[:, 0] = d8[:, -1] = ND
    return d8


@pytest.mark.parametrize("seed", range(8))
def test_converge_matches_brute_force(seed):
    rng = np.random.default_rng(100 + seed)
    d8 = acyclic_codes(rng, 12, 15, nodata_frac=0.05 if seed % 2 else 0.0)
    full, _ = po.sca_d8(d8, converge=True)
    ref = brute_upstream(d8)
    # brute force counts inflows from interior NoData-free paths only; the
    # last-cell rule is off in the converging variant
    assert np.array_equal(full, ref)


@pytest.mark.parametrize("seed", range(12))
def test_one_sweep_matches_recursive_model(seed):
    rng = np.random.default_rng(200 + seed)
    rows, cols = int(rng.integers(3, 14)), int(rng.integers(3, 14))
    if seed % 3 == 0:
        d8 = acyclic_codes(rng, rows, cols, nodata_frac=0.1)
    else:
        # arbitrary codes, cycles, mutual pairs and codes on the border
        d8 = rng.choice([-1.0, 0, 1, 2, 3, 4, 5, 6, 7, 8, ND], size=(rows, cols))
    s, sweeps = po.sca_d8(d8)
    assert np.array_equal(s, model_one_sweep(d8)), (rows, cols)


def test_mutual_pair_is_excluded():
    # (1,1) drains E into (1,2), which drains W back: neither counts the other
    inner = np.array([[5.0, 3.0, 7.0], [-1.0, -1.0, -1.0]])
    d8 = bordered(inner)
    s, _ = po.sca_d8(d8)
    assert s[1, 1] == 1 and s[1, 2] == 1
    assert s[2, 3] == 2  # (1,3) drains S into (2,3)


def test_border_inflow_never_arrives():
    # a border cell draining into the interior keeps its target waiting
    d8 = bordered(np.array([[7.0, 7.0], [-1.0, -1.0]]))
    d8[0, 1] = 7.0  # border cell above (1,1) drains south
    s, _ = po.sca_d8(d8)
    full, _ = po.sca_d8(d8, converge=True)
    assert s[1, 1] == 1 and s[2, 1] == 1  # (1,1) never finishes, so (2,1) gets nothing
    assert s[2, 2] == 2
    assert full[2, 1] == 1 and full[2, 2] == 2


def test_last_cell_nodata_adds_inflows():
    # 3x3 interior; the last work cell (3,3) is NoData and its NW and N
    # neighbours drain into it
    inner = np.array([[-1.0, -1.0, -1.0], [-1.0, 8.0, 7.0], [-1.0, -1.0, ND]])
    d8 = bordered(inner)
    s, _ = po.sca_d8(d8)
    assert s[3, 3] == ND + 2
    full, _ = po.sca_d8(d8, converge=True)
    assert full[3, 3] == ND


def test_tiny_grids_are_all_nodata():
    for shape in [(1, 1), (2, 5), (5, 2)]:
        s, _ = po.sca_d8(np.full(shape, 7.0))
        assert np.all(s == ND)


def test_nan_cells_take_no_inflow():
    inner = np.array([[7.0, -1.0], [np.nan, -1.0], [-1.0, -1.0]])
    d8 = bordered(inner)
    s, _ = po.sca_d8(d8)
    assert s[2, 1] == 1  # NaN: area 1, the inflow from (1,1) is lost
    assert s[3, 1] == 1


def test_random_codes_many_sizes_match_model():
    rng = random.Random(7)
    for _ in range(40):
        rows, cols = rng.randint(1, 9), rng.randint(1, 9)
        d8 = np.array([[rng.choice([-1.0, 1, 3, 5, 7, 0, 2, 6, 8, ND]) for _ in range(cols)] for _ in range(rows)])
        s, _ = po.sca_d8(d8)
        assert np.array_equal(s, model_one_sweep(d8))


def model_fixed_point(d8, nd=ND):
    """Converged accumulation as a least fixed point: a cell is done once
    every inflow is an interior source or done; its area sums the inflows
    that are done (or sources)."""
    rows, cols = d8.shape
    interior = lambda r, c: 1 <= r <= rows - 2 and 1 <= c <= cols - 2
    valid = lambda r, c: abs(d8[r, c] - nd) > 1e-6
    nodata = lambda r, c: abs(d8[r, c] - nd) < 1e-6
    ins = {}
    for r in range(1, rows - 1):
        for c in range(1, cols - 1):
            if valid(r, c):
                ins[(r, c)] = [(r + dr, c + dc) for k, (dr, dc) in enumerate(OFF)
                               if k != 4 and d8[r + dr, c + dc] == 8 - k and d8[r, c] + (8 - k) != 8]
    done = {c: not v for c, v in ins.items()}
    val = {c: 1.0 for c in ins}
    changed = True
    while changed:
        changed = False
        for c, us in ins.items():
            if done[c]:
                continue
            avail = [u for u in us if interior(*u) and u in done and done[u]]
            val[c] = 1.0 + sum(val[u] for u in avail)
            if len(avail) == len(us):
                done[c] = changed = True
    out = np.full(d8.shape, nd)
    for r in range(1, rows - 1):
        for c in range(1, cols - 1):
            if not nodata(r, c):
                out[r, c] = val.get((r, c), 1.0)
    return out


@pytest.mark.parametrize("seed", range(10))
def test_converge_matches_fixed_point_model(seed):
    rng = np.random.default_rng(300 + seed)
    rows, cols = int(rng.integers(3, 13)), int(rng.integers(3, 13))
    d8 = rng.choice([-1.0, 0, 1, 2, 3, 4, 5, 6, 7, 8, ND, np.nan], size=(rows, cols))
    full, _ = po.sca_d8(d8, converge=True)
    assert np.array_equal(full, model_fixed_point(d8)), (rows, cols)
