"""Parity at the full sizes of the benchmark workloads, against the oracle
and through size-independent properties:

- slope+D8 on the 16384^2 recipe DEM: D8 codes and the D8 NoData set bit for
  bit, the slope NoData set equal, slope within 1e-5 (max/RMS printed);
- IDW of the 20,000-sample recipe onto 4096^2 cells against the oracle on
  every cell, and its adjoint dL/d(values) for an L2 loss against a seeded
  target grid by the dot-product identity <g, A d> = <A^T g, d> with A the
  oracle's FP64 IDW operator (IDW is linear in the values);
- the slope adjoint on the 8192^2 recipe DEM: FP64 central differences
  (h = 1e-3 m) on 256 random cells, and 50 gradient-descent steps;
- 512 uneven tiles: the graph-captured persistent launch equals per-tile
  launches bit for bit, and every 32nd tile matches the oracle.

Each case prints what it checked, so a failure names its workload.
"""
import os

import numpy as np
import pytest

from oracle import pyoracle
from pargo_amd import autograd, batch, focal, synth
from pargo_amd.idw import IDWOperator

pytestmark = pytest.mark.gpu

CS = 30.0
ND = -9999.0


def nthreads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def test_slope_d8_16384_every_cell():
    import torch
    z = synth.workload_dem("slope_d8")
    print("slope+D8 16384^2: NoData fraction %.4f" % float((z == ND).mean()), flush=True)
    dem = torch.from_numpy(z).cuda()
    s, c, m = focal.slope_d8(dem, CS, ND, want_mask=True)
    torch.cuda.synchronize()
    gs, gc = s.cpu().numpy(), c.cpu().numpy()
    gm = focal.unpack_mask(m, z.shape[1]).cpu().numpy()
    del dem, s, c, m
    rd = pyoracle.d8(z, ND, 256, nthreads=nthreads())
    nd8 = rd == ND
    assert np.array_equal(gm, nd8), "D8 NoData set differs"
    code = np.where(rd == -1, 0, np.where(nd8, 0, rd)).astype(np.uint8)
    del rd
    bad = np.count_nonzero(gc != code)
    assert bad == 0, "%d D8 codes differ" % bad
    rs = pyoracle.slope(z, CS, ND, "FD", nthreads=nthreads())
    nds = rs == ND
    assert np.array_equal(gs == np.float32(ND), nds), "slope NoData set differs"
    v = ~nds
    err = np.abs(gs[v].astype(np.float64) - rs[v])
    rel = err / np.abs(rs[v]).clip(1e-30)
    print("slope 16384^2: %d valid cells, max rel %.3e, RMS rel %.3e" % (v.sum(), rel.max(), np.sqrt(np.mean(rel ** 2))))
    assert np.all(err <= 1e-5 * np.abs(rs[v]) + 1e-12)


@pytest.fixture(scope="module")
def idw_case():
    xs, ys, vs = synth.samples(synth.BASE_SEED + 2, 20000)
    op = IDWOperator(CS, 12, 2, 5000.0, 5000.0).read_samples(xs, ys, vs)
    assert (op.rows, op.cols) == (4096, 4096)
    yield xs, ys, vs, op
    op.free()


def test_idw_20000_samples_4096_every_cell(idw_case):
    import torch
    xs, ys, vs, op = idw_case
    out = op.run()
    torch.cuda.synchronize()
    got = out.cpu().numpy().astype(np.float64)
    ref = pyoracle.idw(xs, ys, vs, CS, k=12, power=2, buffer=5000.0, block_size=5000.0, nthreads=nthreads())
    err = np.abs(got - ref)
    rel = err / np.abs(ref).clip(1e-30)
    print("IDW 4096^2, 20000 samples: max rel %.3e, RMS rel %.3e" % (rel.max(), np.sqrt(np.mean(rel ** 2))))
    assert np.all(err <= 1e-5 * np.abs(ref) + 1e-9)


def test_idw_adjoint_20000_samples_dot_product(idw_case):
    """dL/dv for L = sum (idw(v) - target)^2: with g = 2 (out - target),
    <g, A d> must equal <dL/dv, d> for any d, A the oracle's FP64 IDW map."""
    import torch
    xs, ys, vs, op = idw_case
    out = op.run()
    gen = torch.Generator(device="cuda").manual_seed(synth.BASE_SEED + 3)
    target = out + torch.randn(out.shape, device="cuda", generator=gen)
    g = (2.0 * (out - target)).contiguous()
    adj = op.adjoint(g).cpu().numpy()
    g64 = g.double().cpu().numpy()
    rng = np.random.default_rng(5)
    deltas = {"random": rng.normal(0.0, 1.0, len(vs)),
              "one-hot at the largest gradient": np.eye(1, len(vs), int(np.argmax(np.abs(adj))))[0]}
    for name, d in deltas.items():
        ad = pyoracle.idw(xs, ys, d, CS, k=12, power=2, buffer=5000.0, block_size=5000.0, nthreads=nthreads())
        lhs = float(np.sum(g64 * ad))
        rhs = float(np.dot(adj, d))
        scale = float(np.sum(np.abs(g64) * np.abs(ad)))
        print("IDW adjoint, %s direction: <g,Ad> %.9e  <A^T g,d> %.9e  scale %.3e" % (name, lhs, rhs, scale))
        assert abs(lhs - rhs) <= 1e-9 * scale + 1e-12


def loss_local(z64, t64, r, c):
    """L restricted to the cells whose 3x3 window holds (r, c), in FP64."""
    r0, r1 = max(1, r - 1), min(z64.shape[0] - 2, r + 1)
    c0, c1 = max(1, c - 1), min(z64.shape[1] - 2, c + 1)
    sub = z64[r0 - 1:r1 + 2, c0 - 1:c1 + 2]
    s = pyoracle.slope(sub, CS, ND, "FD")[1:-1, 1:-1]
    t = t64[r0:r1 + 1, c0:c1 + 1]
    keep = s != ND
    return float(np.sum((s[keep] - t[keep]) ** 2))


def test_slope_adjoint_8192_fd_and_descent():
    import torch
    N = 8192
    seed = synth.BASE_SEED + 3
    z = synth.dem(seed, N, N)
    synth.add_nodata_discs(seed, z, fraction=0.01)
    rng = np.random.default_rng(seed)
    zn = (z.astype(np.float64) + rng.normal(0.0, 1.0, z.shape)).astype(np.float32)
    zn[z == ND] = ND
    dem = torch.from_numpy(z).cuda()
    t = focal.slope(torch.from_numpy(zn).cuda(), CS, ND)
    t = torch.where(t == ND, torch.zeros_like(t), t)
    del zn
    # gradient of L = sum_valid (slope(z) - t)^2 through the hand-written adjoint
    d = dem.clone().requires_grad_(True)
    s = autograd.slope(d, CS, ND)
    loss = (((s - t) ** 2) * (s != ND)).sum()
    loss.backward()
    torch.cuda.synchronize()
    g = d.grad.cpu().numpy().astype(np.float64)
    t64 = t.cpu().numpy().astype(np.float64)
    z64 = z.astype(np.float64)
    h = 1e-3
    fd, got = [], []
    while len(fd) < 256:
        r, c = int(rng.integers(1, N - 1)), int(rng.integers(1, N - 1))
        if z[r, c] == ND:
            continue
        lo, hi = max(0, r - 3), min(N, r + 4)
        lc, hc = max(0, c - 3), min(N, c + 4)
        sub = z64[lo:hi, lc:hc].copy()
        tsub = t64[lo:hi, lc:hc]
        rr, cc = r - lo, c - lc
        sub[rr, cc] += h
        lp = loss_local(sub, tsub, rr, cc)
        sub[rr, cc] -= 2 * h
        lm = loss_local(sub, tsub, rr, cc)
        fd.append((lp - lm) / (2 * h))
        got.append(g[r, c])
    fd, got = np.array(fd), np.array(got)
    err = np.abs(got - fd)
    print("slope adjoint 8192^2 vs FP64 FD on 256 cells: max rel %.3e"
          % (err / np.maximum(np.abs(fd), 1e-12)).max())
    assert np.all(err <= 1e-4 * np.abs(fd) + 1e-6 * np.abs(fd).max())
    # 50 gradient-descent steps on the DEM
    losses = []
    cur = dem
    for _ in range(50):
        d = cur.clone().requires_grad_(True)
        s = autograd.slope(d, CS, ND)
        loss = (((s - t) ** 2) * (s != ND)).sum(dtype=torch.float64)
        loss.backward()
        losses.append(loss.item())
        with torch.no_grad():
            cur = torch.where(cur == ND, cur, cur - 200.0 * d.grad)
    print("gradient descent 8192^2: loss %.6e -> %.6e over 50 steps" % (losses[0], losses[-1]))
    assert max(losses) <= losses[0]
    assert losses[-1] < 0.5 * losses[0], losses[::10]


def test_batch_512_tiles_graph_equals_per_tile():
    import torch
    shapes = batch.tile_shapes(512, synth.BASE_SEED + 1000)
    host = [batch.tile_dem(t, r, c) for t, (r, c) in enumerate(shapes)]
    cells = sum(z.size for z in host)
    print("batch: 512 tiles, %.3f Gcells, %.2f GiB of float32 DEM" % (cells / 1e9, 4 * cells / 2 ** 30), flush=True)
    dems = [torch.from_numpy(z).cuda() for z in host]
    b = batch.Batch(dems)
    b.run(batch.PER_TILE)
    torch.cuda.synchronize()
    ref_s = [x.clone() for x in b.slopes]
    ref_c = [x.clone() for x in b.codes]
    b.autotune()
    for x in b.slopes:
        x.fill_(float("nan"))
    for x in b.codes:
        x.fill_(77)
    b.run(batch.GRAPH)
    torch.cuda.synchronize()
    for i in range(len(dems)):
        assert torch.equal(b.codes[i], ref_c[i]), "tile %d D8 graph vs per-tile" % i
        assert torch.equal(b.slopes[i], ref_s[i]), "tile %d slope graph vs per-tile" % i
    for t in range(0, len(dems), 32):
        z = host[t]
        rd = pyoracle.d8(z, ND, 256, nthreads=nthreads())
        nd = rd == ND
        code = np.where(rd == -1, 0, np.where(nd, 0, rd)).astype(np.uint8)
        assert np.array_equal(b.codes[t].cpu().numpy(), code), "tile %d D8 vs oracle" % t
        rs = pyoracle.slope(z, CS, ND, nthreads=nthreads())
        gs = b.slopes[t].cpu().numpy()
        v = rs != ND
        assert np.array_equal(gs == np.float32(ND), ~v), "tile %d slope NoData set" % t
        assert np.all(np.abs(gs[v] - rs[v]) <= 1e-5 * np.abs(rs[v]) + 1e-12), "tile %d slope" % t
    print("batch: graph == per-tile on all 512 tiles; 16 tiles match the oracle")
