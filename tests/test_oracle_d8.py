"""Pins for the oracle's D8 restatement (oracle/pargo_oracle.c, after
apps/hydrology/flowdird8/d8Operator.cpp:28-102).

Known-answer 3x3 windows derived by hand from the reference's rule, the 1e-6
epsilon boundaries placed with exactly representable doubles, and a
rotation property: turning the DEM by 90 degrees turns every ESRI code by
two bit positions, which fails for any swapped neighbour index or weight.
"""
import numpy as np
import pytest

from oracle import pyoracle

ND = -9999.0
# window index -> (dr, dc) and ESRI code (d8Operator.cpp:96-97)
WIN = {0: (-1, -1), 1: (-1, 0), 2: (-1, 1), 3: (0, -1), 5: (0, 1), 6: (1, -1), 7: (1, 0), 8: (1, 1)}
ESRI = {5: 1, 8: 2, 7: 4, 6: 8, 3: 16, 0: 32, 1: 64, 2: 128}


def window(center, nbrs, default=None):
    """3x3 DEM padded to 5x5 so the centre cell is interior."""
    d = center if default is None else default
    z = np.full((5, 5), d, dtype=np.float64)
    z[2, 2] = center
    for w, v in nbrs.items():
        dr, dc = WIN[w]
        z[2 + dr, 2 + dc] = v
    return z


def code_at(z, dirtype=256):
    return pyoracle.d8(z, ND, dirtype)[2, 2]


@pytest.mark.parametrize("w", sorted(WIN))
def test_single_steepest_direction(w):
    z = window(10.0, {w: 9.0}, default=10.0)
    assert code_at(z, 256) == ESRI[w]
    assert code_at(z, 8) == w


def test_diagonal_needs_sqrt2_more_drop():
    # cardinal drop 1.0 against diagonal raw drop 1.5 (1.06 after /sqrt 2)
    z = window(10.0, {5: 9.0, 8: 8.5}, default=10.0)
    assert code_at(z) == ESRI[8]
    # diagonal raw drop 1.4 (0.99 after /sqrt 2) loses to the cardinal 1.0
    z = window(10.0, {5: 9.0, 8: 8.6}, default=10.0)
    assert code_at(z) == ESRI[5]


def test_cardinal_diagonal_tie_within_epsilon():
    # float32(sqrt 2) / sqrt(2.0) = 1 - 3e-8: within 1e-6 of the cardinal 1.0
    f = float(np.float32(np.sqrt(2.0)))
    z = window(10.0, {5: 9.0, 8: 10.0 - f}, default=10.0)
    assert code_at(z) == ESRI[5] + ESRI[8]


def test_exact_ties_sum_codes():
    z = window(10.0, {1: 8.0, 3: 8.0, 7: 8.0}, default=10.0)
    assert code_at(z, 256) == ESRI[1] + ESRI[3] + ESRI[7]


@pytest.mark.parametrize("tied,median", [([1, 5], 5), ([1, 3, 5], 3), ([0, 2, 6, 8], 6), ([1, 3, 5, 7], 5),
                                         ([3], 3), ([1, 3, 5, 7, 0, 2, 6, 8], 5)])
def test_dirtype8_median_rule(tied, median):
    # validDir lists tied window indices in ascending order; the output is
    # validDir[len >> 1].  Diagonals drop sqrt(2) as much to tie exactly.
    nb = {}
    for w in tied:
        nb[w] = 10.0 - (np.sqrt(2.0) if w in (0, 2, 6, 8) else 1.0)
    z = window(10.0, nb, default=10.0)
    assert code_at(z, 8) == median
    assert code_at(z, 256) == sum(ESRI[w] for w in tied)


def test_tie_epsilon_boundaries():
    # drops 1 and 1 - 0.5e-6: tie; 1 and 1 - 2e-6: no tie
    z = window(2.0, {5: 1.0, 7: 1.0 + 0.5e-6}, default=2.0)
    assert code_at(z) == ESRI[5] + ESRI[7]
    z = window(2.0, {5: 1.0, 7: 1.0 + 2e-6}, default=2.0)
    assert code_at(z) == ESRI[5]


def test_flat_threshold():
    z = window(1.0, {3: 1.0 - 0.9e-6}, default=1.0)
    assert code_at(z) == -1
    z = window(1.0, {3: 1.0 - 1.1e-6}, default=1.0)
    assert code_at(z) == ESRI[3]


def test_flat_and_pit_are_minus_one():
    assert code_at(np.full((5, 5), 4.0)) == -1
    assert code_at(window(1.0, {}, default=5.0)) == -1
    assert code_at(window(1.0, {}, default=5.0), 8) == -1


def test_nodata_window_and_epsilon():
    z = window(10.0, {6: ND}, default=10.0)
    assert code_at(z) == ND
    z = window(10.0, {6: ND + 0.5e-6}, default=10.0)  # within 1e-6
    assert code_at(z) == ND
    z = window(10.0, {6: ND + 2e-6}, default=10.0)  # outside: a huge drop
    assert code_at(z) == ESRI[6]


def test_border_is_nodata():
    rng = np.random.default_rng(0)
    d = pyoracle.d8(rng.uniform(0, 10, size=(6, 7)), ND, 256)
    assert np.all(d[[0, -1], :] == ND) and np.all(d[:, [0, -1]] == ND)
    assert np.all(d[1:-1, 1:-1] != ND)


def _rot_code(c, quarter_turns_ccw):
    if c < 0:
        return c
    c = int(c)
    s = (6 * quarter_turns_ccw) % 8  # E(bit0) -> N(bit6) per turn
    return ((c << s) | (c >> (8 - s))) & 0xFF if s else c


def test_rotation_equivariance():
    rng = np.random.default_rng(4)
    z = rng.integers(0, 6, size=(40, 40)).astype(np.float64) * 3.0
    z[rng.random(z.shape) < 0.3] += rng.uniform(0, 1, size=z.shape)[rng.random(z.shape) < 0.3].mean()
    base = pyoracle.d8(z, ND, 256)
    for k in (1, 2, 3):
        rot = pyoracle.d8(np.rot90(z, k).copy(), ND, 256)
        expect = np.vectorize(lambda c: _rot_code(c, k))(np.rot90(base, k))
        assert np.array_equal(rot, expect)


def test_dirtype8_consistent_with_256():
    """dirType 8 = median tied window index of the tie set the 256 code
    lists (checked on random terrain with many ties)."""
    rng = np.random.default_rng(5)
    z = rng.integers(0, 4, size=(60, 70)).astype(np.float64)
    c256 = pyoracle.d8(z, ND, 256)
    c8 = pyoracle.d8(z, ND, 8)
    bit_to_win = {v: k for k, v in ESRI.items()}
    for v256, v8 in zip(c256.ravel(), c8.ravel()):
        if v256 in (ND, -1):
            assert v8 == v256
            continue
        wins = sorted(bit_to_win[1 << b] for b in range(8) if int(v256) >> b & 1)
        assert v8 == wins[len(wins) >> 1]
