"""GPU parity of the compute-load model against the oracle.

  intensity ratio  pargo_intensity_load vs oracle_intensity_load (bit-exact:
                   sums of the same doubles in the same order)
  experiment       pargo_idw_work: per-cell samples examined; per-row sums
                   equal the oracle's per-row candidate counts, masked cells
                   0, and the values of the same run equal pargo_idw_run's
"""
import numpy as np
import pytest

from oracle import pyoracle
from pargo_amd import _lib, compute, synth
from pargo_amd.idw import IDWOperator

pytestmark = pytest.mark.gpu

ND = -9999.0


def dev(a, pitch_pad=0):
    import torch
    rows, cols = a.shape
    buf = torch.zeros((rows, cols + pitch_pad), dtype=torch.float64, device="cuda")
    buf[:, :cols] = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    return buf[:, :cols]


@pytest.mark.parametrize("seed", range(5))
def test_intensity_load_matches_oracle(seed):
    rng = np.random.default_rng(700 + seed)
    rows, cols = int(rng.integers(3, 300)), int(rng.integers(3, 300))
    nl = 1 + seed % 3
    layers, nds = [], []
    for _ in range(nl):
        a = rng.random((rows, cols)) * 100
        nd = ND if rng.random() < 0.5 else -1.0
        a[rng.random((rows, cols)) < 0.2] = nd
        a[rng.random((rows, cols)) < 0.02] = nd + 5e-8   # NoData by the 1e-7 test
        a[rng.random((rows, cols)) < 0.02] = nd + 3e-7   # valid
        layers.append(a)
        nds.append(nd)
    g = int(rng.integers(1, 12))
    r0, c0 = (1, 1) if seed % 2 == 0 else (0, 0)
    wr, wc = rows - 2 * r0, cols - 2 * c0
    nodata_load, valid_load = float(rng.random()), 1.0 + float(rng.random())
    ref = pyoracle.intensity_load(layers, nds, r0, c0, wr, wc, g, nodata_load, valid_load)
    got = compute.intensity_load([dev(a, seed % 2) for a in layers], nds, g, nodata_load, valid_load,
                                 (r0, c0, wr, wc)).cpu().numpy()
    assert np.array_equal(got, ref)


def test_intensity_load_refuses_negative_loads():
    with pytest.raises(_lib.PargoError):
        compute.intensity_load([dev(np.ones((8, 8)))], [ND], 2, -1.0, 1.0)
    with pytest.raises(ValueError):
        compute.Transformation(0.0, -2.0)


def test_idw_work_matches_oracle_candidates():
    import torch
    xs, ys, vs = synth.samples(synth.BASE_SEED + 9, 3000, width=30000.0, height=24000.0)
    op = IDWOperator(100.0, 8, 2, 3000.0, 2500.0).read_samples(xs, ys, vs)
    out = op.run()
    vals = torch.empty_like(out)
    work = op.work(out=vals)
    torch.cuda.synchronize()
    assert torch.equal(vals, out)
    _, cand, _ = pyoracle.idw(xs, ys, vs, 100.0, k=8, power=2, buffer=3000.0, block_size=2500.0, stats=True)
    w = work.cpu().numpy()
    assert np.array_equal(w.sum(axis=1).astype(np.int64), cand)
    # masked cells examine nothing
    m = torch.ones((op.rows, op.cols), dtype=torch.int32, device="cuda")
    m[: op.rows // 2] = 0
    op.mask_layer(m, -1)
    wm = op.work().cpu().numpy()
    assert np.all(wm[: op.rows // 2] == 0) and np.array_equal(wm[op.rows // 2:], w[op.rows // 2:])
    op.free()
