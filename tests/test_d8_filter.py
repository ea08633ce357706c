"""The FP32 D8 filter of pargo_amd/csrc/focal.hip (focal_step), restated in
numpy float32 arithmetic, never claims a code that differs from the
reference's FP64 decision (d8Operator.cpp:61-97, via the oracle).

The kernel settles a cell in FP32 when the filter proves its code, and
recomputes the lane in FP64 otherwise; so the filter may call any cell
ambiguous, but a cell it settles must carry the oracle's code.  The
inputs are chosen to sit on the filter's edges: drops of a few 2^-20
around the 1e-6 flat and tie tests, cardinal/diagonal near-ties, exact
integer ties, tiny and large elevations, NaN.  A filter that dropped the
tie margin, used the wrong sign for a negated drop or tested the flat
bound on the wrong side fails here.
"""
import numpy as np
import pytest

from oracle import pyoracle

INV_SQRT2_F = np.float32(float.fromhex("0x1.6a09e6p-1"))
K = np.float32(1.0001e-6)
SHRINK = np.float32(float.fromhex("0x1.ffffep-1"))
FLOOR = np.float32(float.fromhex("0x1.0c6898p-20"))
BAND_MID = np.float32(1.5e-6)
BAND_HALF = np.float32(0.50025e-6)
# window offsets in ESRI bit order (E=1, SE=2, S=4, SW=8, W=16, NW=32, N=64, NE=128)
ESRI = [(0, 1), (1, 1), (1, 0), (1, -1), (0, -1), (-1, -1), (-1, 0), (-1, 1)]


def fp32_filter(z):
    """(code, ambiguous) for interior cells, as the kernel computes them."""
    z = np.asarray(z, dtype=np.float32)
    c = z[1:-1, 1:-1]
    drops = []
    with np.errstate(invalid="ignore", over="ignore"):
        for dr, dc in ESRI:
            nb = z[1 + dr:z.shape[0] - 1 + dr, 1 + dc:z.shape[1] - 1 + dc]
            d = (c - nb).astype(np.float32)
            if dr and dc:
                d = (d * INV_SQRT2_F).astype(np.float32)
            drops.append(d)
        D = np.stack(drops)
        M = np.maximum(np.nanmax(np.where(np.isnan(D), -np.inf, D), axis=0), np.float32(0)).astype(np.float32)
        # fma(M, 1 - 2^-20, -K) rounded once: the product and difference are
        # exact in long double for these magnitudes
        L = (M.astype(np.longdouble) * np.longdouble(SHRINK) - np.longdouble(K)).astype(np.float32)
        L = np.maximum(L, FLOOR)
        cand = D > L[None]
    code = np.zeros(c.shape, np.int64)
    for b in range(8):
        code |= cand[b].astype(np.int64) << b
    pop = cand.sum(axis=0)
    band = np.abs((M - BAND_MID).astype(np.float32)) < BAND_HALF
    return code, (pop > 1) | band


def oracle_codes(z):
    ref = pyoracle.d8(z, nodata=-1e30, dirtype=256)[1:-1, 1:-1]
    return np.where(ref == -1, 0, ref).astype(np.int64)


def check(z, max_amb=None):
    code, amb = fp32_filter(z)
    ref = oracle_codes(z)
    bad = np.argwhere(~amb & (code != ref))
    assert bad.size == 0, "filter settles %s with %s, oracle %s" % (
        bad[:5].tolist(), code[tuple(bad[:5].T)], ref[tuple(bad[:5].T)])
    if max_amb is not None:
        assert amb.mean() <= max_amb, amb.mean()
    return amb.mean()


@pytest.mark.parametrize("base", [0.0, 1.0, 3.0, 1000.0])
def test_units_of_2_pow_20(base):
    rng = np.random.default_rng(int(base) + 1)
    for step in (2.0 ** -20, 2.0 ** -21, 3 * 2.0 ** -22):
        z = (base + rng.integers(0, 6, size=(120, 150)) * step).astype(np.float32)
        check(z)


def test_drops_next_to_the_flat_bound():
    # the centre sits 0.9999e-6 .. 2.1e-6 above all neighbours, float by float
    rng = np.random.default_rng(2)
    lo = np.float32(0.999e-6)
    n = 4000
    steps = rng.integers(0, 9000, size=n)
    vals = (lo + steps.astype(np.float64) * 1.3e-10).astype(np.float32)
    z = np.zeros((3, 3 * n), np.float32)
    z[1, 1::3] = vals
    check(z)


def test_cardinal_diagonal_near_ties():
    rng = np.random.default_rng(3)
    h = rng.uniform(1e-6, 5, size=(160, 200))
    z = np.zeros((160, 200), dtype=np.float32)
    z[::2, ::2] = 1000.0
    z[1::2, ::2] = (1000.0 - h[1::2, ::2]).astype(np.float32)
    z[::2, 1::2] = (1000.0 - np.sqrt(2) * h[::2, 1::2]).astype(np.float32)
    z[1::2, 1::2] = (1000.0 - np.sqrt(2) * h[1::2, 1::2]).astype(np.float32)
    check(z)


def test_ties_within_a_few_ulps():
    # neighbours at the same drop, then nudged by 1..3 float ulps
    rng = np.random.default_rng(4)
    z = np.full((150, 150), 500.0, np.float32)
    z[::2, ::2] = 510.0
    nudge = rng.integers(-3, 4, size=z.shape).astype(np.float32)
    z = np.nextafter(z, z + np.sign(nudge)).astype(np.float32)
    z[rng.random(z.shape) < 0.3] = 500.0
    check(z)


def test_integer_ties_and_plateaus():
    rng = np.random.default_rng(5)
    z = (rng.integers(0, 4, size=(200, 220)) * 10).astype(np.float32)
    check(z)


@pytest.mark.parametrize("scale", [1e-5, 1e-3, 1.0, 1e4, 1e7])
def test_random_terrain_scales(scale):
    rng = np.random.default_rng(int(np.log10(scale)) + 20)
    z = np.cumsum(np.cumsum(rng.normal(size=(180, 180)), axis=0), axis=1)
    z = (z / np.abs(z).max() * scale).astype(np.float32)
    check(z)


def test_negative_and_mixed_sign():
    rng = np.random.default_rng(6)
    z = rng.uniform(-3e-6, 3e-6, size=(200, 200)).astype(np.float32)
    check(z)
    z = rng.uniform(-50, 50, size=(200, 200)).astype(np.float32)
    check(z)


def test_nan_windows():
    rng = np.random.default_rng(7)
    z = rng.uniform(0, 10, size=(120, 120)).astype(np.float32)
    z[rng.random(z.shape) < 0.05] = np.nan
    check(z)


def test_terrain_is_mostly_settled():
    from pargo_amd import synth
    z = synth.dem(11, 256, 512)
    synth.add_plateaus(11, z, 0.02)
    rate = check(z, max_amb=0.002)
    print("ambiguous fraction %.5f" % rate)
