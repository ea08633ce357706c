"""GPU parity of the D8 contributing area (scad8) against the oracle.

Bar: bit-exact (areas are integer counts held in float64; NoData cells and
the reference's last-cell value included), for both the reference's single
accumulation sweep and the converging variant.  At 16384^2 the converging
result is checked through the recurrence it must satisfy exactly:
area(c) = 1 + sum of the areas of the cells draining into c.
"""
import numpy as np
import pytest

from oracle import pyoracle
from pargo_amd import _lib, focal, hydro, synth

pytestmark = pytest.mark.gpu

ND = -9999.0
OFF = [(-1, -1), (-1, 0), (-1, 1), (0, -1), (0, 0), (0, 1), (1, -1), (1, 0), (1, 1)]


def gpu_from_double(d8, nodata=ND, mode="reference", pitch_pad=0):
    import torch
    rows, cols = d8.shape
    buf = torch.zeros((rows, cols + pitch_pad), dtype=torch.float64, device="cuda")
    buf[:, :cols] = torch.from_numpy(np.ascontiguousarray(d8, dtype=np.float64)).cuda()
    t = buf[:, :cols]
    flow = hydro.d8_classify(t, nodata)
    out = hydro.sca_d8(flow, nodata, mode=mode)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def d8_of_dem(z, nodata=ND):
    """pargo_d8 dirtype 8 on the device: (codes, mask words, reference layer)."""
    import torch
    t = torch.from_numpy(np.ascontiguousarray(z, dtype=np.float32)).cuda()
    code, mask = focal.d8(t, 30.0, nodata, dirtype=8, want_mask=True)
    ref = focal.codes_to_reference(code, focal.unpack_mask(mask, z.shape[1]), 8, nodata)
    return code, mask, ref


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("mode", ["reference", "converge"])
def test_random_codes_match_oracle(seed, mode):
    rng = np.random.default_rng(1000 + seed)
    rows, cols = int(rng.integers(1, 70)), int(rng.integers(1, 300))
    choice = [-1.0, 0, 1, 2, 3, 4, 5, 6, 7, 8, ND, np.nan, 2.9999999999999996, 5.5]
    d8 = rng.choice(choice, size=(rows, cols))
    if seed % 2:
        d8[0, :] = d8[-1, :] = d8[:, 0] = d8[:, -1] = ND
    ref, _ = pyoracle.sca_d8(d8, ND, mode == "converge")
    got = gpu_from_double(d8, ND, mode, pitch_pad=seed % 3)
    assert np.array_equal(got, ref), np.argwhere(got != ref)[:5].tolist()


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("mode", ["reference", "converge"])
def test_random_codes_interior_tiles_match_oracle(seed, mode):
    """Rasters large enough for the init pass's word-wise interior tiles
    (64 x 32 cells, 2+ cells from the border, class rows read as aligned
    words), every class kind, row pitches at every byte offset."""
    rng = np.random.default_rng(2000 + seed)
    rows, cols = 200 + int(rng.integers(0, 40)), 400 + int(rng.integers(0, 80))
    choice = [-1.0, 0, 1, 2, 3, 4, 5, 6, 7, 8, ND, np.nan, 2.9999999999999996, 5.5]
    p = np.array([4, 8, 8, 8, 2, 8, 8, 8, 8, 1, 1, 1, 1, 1], dtype=np.float64)
    d8 = rng.choice(choice, size=(rows, cols), p=p / p.sum())
    ref, _ = pyoracle.sca_d8(d8, ND, mode == "converge")
    got = gpu_from_double(d8, ND, mode, pitch_pad=seed)
    assert np.array_equal(got, ref), np.argwhere(got != ref)[:5].tolist()


@pytest.mark.parametrize("mode", ["reference", "converge"])
def test_dem_d8_chain_matches_oracle(mode):
    import torch
    z = synth.dem(synth.BASE_SEED + 50, 700, 513)
    z = synth.add_plateaus(synth.BASE_SEED + 51, z, 0.02)
    synth.add_nodata_discs(synth.BASE_SEED + 52, z, n_discs=8)  # in place
    code, mask, ref_layer = d8_of_dem(z)
    d8 = ref_layer.cpu().numpy()
    # the device codes equal the oracle's D8 layer
    assert np.array_equal(d8, pyoracle.d8(z, ND, 8))
    ref, _ = pyoracle.sca_d8(d8, ND, mode == "converge")
    got_codes = hydro.sca_d8(code, ND, mask, mode).cpu().numpy()
    got_double = gpu_from_double(d8, ND, mode)
    torch.cuda.synchronize()
    assert np.array_equal(got_codes, ref)
    assert np.array_equal(got_double, ref)


def test_last_cell_rule_on_device():
    inner = np.array([[-1.0, -1.0, -1.0], [-1.0, 8.0, 7.0], [-1.0, -1.0, ND]])
    d8 = np.full((5, 5), ND)
    d8[1:-1, 1:-1] = inner
    got = gpu_from_double(d8)
    assert got[3, 3] == ND + 2
    assert gpu_from_double(d8, mode="converge")[3, 3] == ND


@pytest.mark.parametrize("nodata", [float("nan"), 0.0, -1.0, 5.0, 8.0000015])
def test_unrepresentable_nodata_is_refused(nodata):
    import torch
    d8 = torch.full((8, 8), 7.0, dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.PargoError):
        hydro.d8_classify(d8, nodata)
    flow = torch.full((8, 8), 7, dtype=torch.uint8, device="cuda")
    with pytest.raises(_lib.PargoError):
        hydro.sca_d8(flow, nodata)


def test_bad_mode_is_refused():
    import ctypes
    import torch
    flow = torch.full((4, 4), 7, dtype=torch.uint8, device="cuda")
    out = torch.empty((4, 4), dtype=torch.float64, device="cuda")
    ctx = _lib.default_context(0)
    g = focal.grid_of(flow, 1.0, ND)
    rc = _lib.lib().pargo_sca_d8(ctx.h, ctypes.byref(g), _lib._ptr(flow), None, 7, _lib._ptr(out))
    assert rc == _lib.PARGO_E_INVALID


def test_full_size_converge_recurrence():
    """16384^2 DEM -> D8 (dirType 8) -> converged area; every valid interior
    cell must equal 1 + the sum of its inflows' areas (all inflows here come
    from valid interior cells or are absent)."""
    import torch
    z = synth.workload_dem("slope_d8")
    code, mask, _ = d8_of_dem(z)
    del z
    area = hydro.sca_d8(code, ND, mask, "converge")
    nd = focal.unpack_mask(mask, code.shape[1])
    rows, cols = code.shape
    c = code.to(torch.int16)
    expect = torch.ones((rows - 2, cols - 2), dtype=torch.float64, device="cuda")
    valid_area = torch.where(nd, torch.zeros_like(area), area)
    for k, (dr, dc) in enumerate(OFF):
        if k == 4:
            continue
        u = c[1 + dr:rows - 1 + dr, 1 + dc:cols - 1 + dc]
        flows_in = (u == 8 - k) & ~nd[1 + dr:rows - 1 + dr, 1 + dc:cols - 1 + dc]
        expect += torch.where(flows_in, valid_area[1 + dr:rows - 1 + dr, 1 + dc:cols - 1 + dc],
                              torch.zeros_like(expect))
    inner_nd = nd[1:-1, 1:-1]
    got = area[1:-1, 1:-1]
    assert torch.equal(torch.where(inner_nd, torch.zeros_like(got), got),
                       torch.where(inner_nd, torch.zeros_like(expect), expect))
    assert torch.all(area[nd] == ND)
    # every valid cell is counted once, at the outlet its path ends in: the
    # areas of cells draining nowhere valid sum to the number of valid cells
    valid = ~nd
    drains_valid = torch.zeros_like(valid)
    for k, (dr, dc) in enumerate(OFF):
        if k == 4:
            continue
        tgt = torch.zeros_like(valid)
        tgt[1:-1, 1:-1] = valid[1 + dr:rows - 1 + dr, 1 + dc:cols - 1 + dc]
        drains_valid |= (code == k) & tgt
    outlets = valid & ~drains_valid
    assert int(area[outlets].sum().item()) == int(valid.sum().item())
