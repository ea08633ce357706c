"""Batched slope+D8 (pargo_batch_slope_d8): every mode gives the same bytes
as per-raster pargo_slope_d8 calls, D8 matches the oracle bit for bit, and
graph replays stay correct with other GPU work in between."""
import numpy as np
import pytest

from oracle import pyoracle
from pargo_amd import batch, focal

pytestmark = pytest.mark.gpu


def make_tiles(n, seed, smin, smax):
    import torch
    shapes = batch.tile_shapes(n, seed, smin, smax)
    return [torch.from_numpy(batch.tile_dem(t, r, c)).cuda() for t, (r, c) in enumerate(shapes)]


def eager(dems):
    import torch
    out = [focal.slope_d8(d, 30.0, -9999.0) for d in dems]
    torch.cuda.synchronize()
    return [(s.clone(), c.clone()) for s, c, _ in out]


@pytest.mark.parametrize("mode", [batch.PER_TILE, batch.PERSISTENT, batch.GRAPH])
def test_batch_modes_match_single_calls(mode):
    import torch
    dems = make_tiles(40, 7, 3, 300)
    ref = eager(dems)
    b = batch.Batch(dems)
    s, c = b.run(mode)
    torch.cuda.synchronize()
    for i in range(len(dems)):
        assert torch.equal(c[i], ref[i][1]), "tile %d D8" % i
        assert torch.equal(s[i], ref[i][0]), "tile %d slope" % i


def test_batch_d8_matches_oracle():
    import torch
    dems = make_tiles(12, 8, 16, 400)
    b = batch.Batch(dems)
    s, c = b.run(batch.GRAPH)
    torch.cuda.synchronize()
    for i, d in enumerate(dems):
        z = d.cpu().numpy()
        rd = pyoracle.d8(z, -9999.0, 256)
        nd = rd == -9999.0
        code = np.where(rd == -1, 0, np.where(nd, 0, rd)).astype(np.int64)
        got = np.where(nd, 0, c[i].cpu().numpy().astype(np.int64))
        assert np.array_equal(got, code), "tile %d" % i
        rs = pyoracle.slope(z, 30.0, -9999.0)
        v = rs != -9999.0
        g = s[i].cpu().numpy().astype(np.float64)
        assert np.all(np.abs(g[v] - rs[v]) <= 1e-5 * np.abs(rs[v]) + 1e-12)


def test_graph_replays_with_eager_work_between():
    import torch
    dems = make_tiles(30, 9, 8, 256)
    ref = eager(dems)
    b = batch.Batch(dems)
    for k in range(4):
        for x in b.slopes:
            x.fill_(float("nan"))
        for x in b.codes:
            x.zero_()
        b.run(batch.GRAPH)
        junk = torch.randn(257, 333, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        for i in range(len(dems)):
            assert torch.equal(b.codes[i], ref[i][1]), "replay %d tile %d" % (k, i)
            assert torch.equal(b.slopes[i], ref[i][0]), "replay %d tile %d" % (k, i)


def test_batch_with_special_values_uses_fixup():
    import torch
    dems = make_tiles(10, 10, 20, 120)
    dems[3][5, 5] = float("inf")
    dems[7][10, 11] = 3e20
    ref = eager(dems)
    b = batch.Batch(dems)
    b.run(batch.PERSISTENT)
    torch.cuda.synchronize()
    for i in range(len(dems)):
        assert torch.equal(b.codes[i], ref[i][1])
        assert torch.equal(torch.nan_to_num(b.slopes[i], nan=-1.0), torch.nan_to_num(ref[i][0], nan=-1.0))


def test_autotune_keeps_results():
    import torch
    dems = make_tiles(20, 11, 64, 700)
    ref = eager(dems)
    b = batch.Batch(dems)
    cpt, rows = b.autotune()
    assert cpt in (2, 4) and rows in (16, 32, 64, 128)
    b.run(batch.GRAPH)
    torch.cuda.synchronize()
    for i in range(len(dems)):
        assert torch.equal(b.codes[i], ref[i][1])
        assert torch.equal(b.slopes[i], ref[i][0])


def test_mixed_cell_sizes_all_modes_agree():
    """A tile whose cell size is outside the Horn fast path's range sends the
    whole batch down the generic FP64 path in every mode: the three modes
    give the same bits, D8 equals single calls, slope stays within 1e-5."""
    import torch
    dems = make_tiles(12, 21, 8, 200)
    sizes = [30.0] * len(dems)
    sizes[4] = 1e-7  # below 2^-20
    single = []
    for d, cs in zip(dems, sizes):
        s_, c_, _ = focal.slope_d8(d, cs, -9999.0)
        single.append((s_.clone(), c_.clone()))
    torch.cuda.synchronize()
    outs = []
    for mode in (batch.PER_TILE, batch.PERSISTENT, batch.GRAPH):
        b = batch.Batch(dems, cellsize=sizes)
        s, c = b.run(mode)
        torch.cuda.synchronize()
        outs.append(([x.clone() for x in s], [x.clone() for x in c]))
    for m in (1, 2):
        for i in range(len(dems)):
            assert torch.equal(outs[m][0][i], outs[0][0][i]), "mode %d tile %d slope" % (m, i)
            assert torch.equal(outs[m][1][i], outs[0][1][i]), "mode %d tile %d D8" % (m, i)
    for i in range(len(dems)):
        assert torch.equal(outs[0][1][i], single[i][1]), "tile %d D8 vs single call" % i
        ref = single[i][0].double()
        got = outs[0][0][i].double()
        assert torch.all((got - ref).abs() <= 1e-5 * ref.abs() + 1e-12), "tile %d slope vs single call" % i


@pytest.mark.parametrize("source", ["rows", "measured", "estimate"])
def test_queue_order_does_not_change_values(source):
    """pargo_batch_order (cells, measured per-tile times, or as given) only
    changes when tiles start: every mode still equals the single calls."""
    import torch
    dems = make_tiles(24, 31, 3, 260)
    ref = eager(dems)
    b = batch.Batch(dems)
    b.order(source)
    for mode in (batch.PER_TILE, batch.PERSISTENT, batch.GRAPH):
        s, c = b.run(mode)
        torch.cuda.synchronize()
        for i in range(len(dems)):
            assert torch.equal(c[i], ref[i][1]), "%s mode %d tile %d D8" % (source, mode, i)
            assert torch.equal(s[i], ref[i][0]), "%s mode %d tile %d slope" % (source, mode, i)
    b.order("estimate")
