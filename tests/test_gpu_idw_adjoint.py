"""IDW adjoint (pargo_idw_adjoint): dL/d(sample values) for
L = sum over cells of (idw(v) - target)^2, against
- FP64 central differences through the oracle (IDW is linear in v, so the
  difference quotient is exact up to rounding), on random samples;
- an FP64 brute-force accumulation of g_c * w_ci / W_c over exact KNN sets
  (buffer 0).
Bar: relative 1e-4 of the largest gradient entry.
"""
import numpy as np
import pytest

from oracle import pyoracle
from pargo_amd import autograd
from pargo_amd.idw import IDWOperator

pytestmark = pytest.mark.gpu


def setup_case(n, seed, extent=4000.0, cs=50.0):
    rng = np.random.default_rng(seed)
    xs = rng.uniform(0, extent, n)
    ys = rng.uniform(0, extent, n)
    vs = rng.uniform(100, 200, n)
    rows, cols, _, _ = pyoracle.idw_dims(xs, ys, cs)
    target = rng.uniform(100, 200, (rows, cols))
    return xs, ys, vs, target, cs


def gpu_grad(xs, ys, vs, target, cs, k, p, buffer, block):
    import torch
    op = IDWOperator(cs, k, p, buffer, block).read_samples(xs, ys, vs)
    v = torch.tensor(vs, dtype=torch.float64, device="cuda", requires_grad=True)
    out = autograd.idw(v, op)
    t = torch.from_numpy(target.astype(np.float32)).cuda()
    loss = ((out.double() - t.double()) ** 2).sum()
    loss.backward()
    torch.cuda.synchronize()
    g = v.grad.cpu().numpy()
    op.free()
    return g


def oracle_loss(xs, ys, vs, target, cs, k, p, buffer, block):
    out = pyoracle.idw(xs, ys, vs, cs, k=k, power=p, buffer=buffer, block_size=block, nthreads=16)
    return float(np.sum((out - target) ** 2))


@pytest.mark.parametrize("buffer", [0.0, 700.0])
def test_idw_adjoint_vs_fd(buffer):
    xs, ys, vs, target, cs = setup_case(200, 1)
    k, p, block = 12, 2, 500.0
    g = gpu_grad(xs, ys, vs, target, cs, k, p, buffer, block)
    rng = np.random.default_rng(2)
    idx = rng.choice(len(vs), 48, replace=False)
    h = 1e-3
    fd = []
    for i in idx:
        vp = vs.copy(); vp[i] += h
        vm = vs.copy(); vm[i] -= h
        fd.append((oracle_loss(xs, ys, vp, target, cs, k, p, buffer, block) -
                   oracle_loss(xs, ys, vm, target, cs, k, p, buffer, block)) / (2 * h))
    fd = np.array(fd)
    err = np.abs(g[idx] - fd)
    print("idw adjoint vs FD: max err %.3e, max |fd| %.3e" % (err.max(), np.abs(fd).max()))
    assert np.all(err <= 1e-4 * np.abs(fd).max())


def test_idw_adjoint_vs_bruteforce_fp64():
    xs, ys, vs, target, cs = setup_case(150, 3, extent=3000.0)
    k, p = 8, 2
    g = gpu_grad(xs, ys, vs, target, cs, k, p, 0.0, 400.0)
    rows, cols, minx, maxy = pyoracle.idw_dims(xs, ys, cs)
    out = pyoracle.idw(xs, ys, vs, cs, k=k, power=p, block_size=400.0)
    gc = 2 * (out - target)
    ref = np.zeros(len(vs))
    csd = float(np.float32(cs))
    for r in range(rows):
        y = maxy - (r + 0.5) * csd
        for c in range(cols):
            x = (c + 0.5) * csd + minx
            d = np.hypot(xs - x, ys - y)
            o = np.argsort(d, kind="stable")[:k]
            w = 1.0 / d[o] ** p
            ref[o] += gc[r, c] * w / w.sum()
    err = np.abs(g - ref)
    print("idw adjoint vs brute force: max err %.3e" % err.max())
    assert err.max() <= 1e-4 * np.abs(ref).max()
