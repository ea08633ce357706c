"""Slope adjoint (pargo_slope_adjoint) against FP64 references.

L(z) = sum over valid cells of (slope(z) - t)^2.
- FP64 central finite differences, step 1e-3 m, on 256 random interior
  cells: relative agreement <= 1e-4 (|gpu - fd| <= 1e-4 * |fd| + 1e-6 * max|fd|);
- an FP64 torch-autograd gradient of the same loss on the CPU (Horn slope
  written with conv2d): max abs error <= 1e-4 * max |grad|.
"""
import numpy as np
import pytest

from oracle import pyoracle
from pargo_amd import autograd, synth

pytestmark = pytest.mark.gpu

CS = 30.0
ND = -9999.0


def dem_for_gradients(rows, cols, seed, nodata_discs=4):
    # no plateaus: a flat window has no derivative of |grad z|
    z = synth.dem(seed, rows, cols)
    if nodata_discs:
        synth.add_nodata_discs(seed, z, n_discs=nodata_discs)
    return z


def torch_loss_grad(z64, t64, cs):
    import torch
    z = torch.tensor(z64, dtype=torch.float64, requires_grad=True)
    kx = torch.tensor([[-1.0, 0.0, 1.0], [-2.0, 0.0, 2.0], [-1.0, 0.0, 1.0]], dtype=torch.float64)
    ky = torch.tensor([[1.0, 2.0, 1.0], [0.0, 0.0, 0.0], [-1.0, -2.0, -1.0]], dtype=torch.float64)
    zz = z[None, None]
    nx = torch.nn.functional.conv2d(zz, kx[None, None])[0, 0]
    ny = torch.nn.functional.conv2d(zz, ky[None, None])[0, 0]
    s = torch.sqrt(nx * nx + ny * ny) / (8 * cs)
    nd = torch.tensor(z64 == ND)
    win = torch.nn.functional.max_pool2d(nd[None, None].double(), 3, stride=1)[0, 0] > 0
    valid = ~win
    t = torch.tensor(t64[1:-1, 1:-1])
    loss = (((s - t) ** 2) * valid).sum()
    loss.backward()
    return z.grad.numpy()


def loss_local(z64, t64, r, c, cs):
    """L restricted to the 3x3 cells whose window holds (r, c), FP64."""
    r0, r1 = max(1, r - 1), min(z64.shape[0] - 2, r + 1)
    c0, c1 = max(1, c - 1), min(z64.shape[1] - 2, c + 1)
    sub = z64[r0 - 1:r1 + 2, c0 - 1:c1 + 2]
    s = pyoracle.slope(sub, cs, ND, "FD")[1:-1, 1:-1]
    t = t64[r0:r1 + 1, c0:c1 + 1]
    m = s != ND
    return float(np.sum((s[m] - t[m]) ** 2))


def gpu_grad(z, t):
    import torch
    dem = torch.from_numpy(z).cuda()
    s = autograd.slope(dem.requires_grad_(True), CS, ND)
    tt = torch.from_numpy(t.astype(np.float32)).cuda()
    valid = s != ND
    loss = (((s - tt) ** 2) * valid).sum()
    loss.backward()
    torch.cuda.synchronize()
    return dem.grad.cpu().numpy().astype(np.float64)


def targets(z, seed):
    rng = np.random.default_rng(seed)
    zn = (z.astype(np.float64) + rng.normal(0, 1.0, z.shape)).astype(np.float32)
    zn[z == ND] = ND
    t = pyoracle.slope(zn, CS, ND, "FD")
    t[t == ND] = 0.0
    return t


@pytest.mark.parametrize("rows,cols", [(200, 260), (67, 1031)])
def test_slope_adjoint_matches_fd(rows, cols):
    z = dem_for_gradients(rows, cols, 21)
    t = targets(z, 22)
    g = gpu_grad(z, t)
    z64 = z.astype(np.float64)
    rng = np.random.default_rng(23)
    h = 1e-3
    checked = 0
    fd = []
    got = []
    while checked < 256:
        r, c = int(rng.integers(1, rows - 1)), int(rng.integers(1, cols - 1))
        if z[r, c] == ND:
            continue
        zp = z64.copy(); zp[r, c] += h
        zm = z64.copy(); zm[r, c] -= h
        fd.append((loss_local(zp, t, r, c, CS) - loss_local(zm, t, r, c, CS)) / (2 * h))
        got.append(g[r, c])
        checked += 1
    fd, got = np.array(fd), np.array(got)
    scale = np.abs(fd).max()
    err = np.abs(got - fd)
    print("adjoint vs FD: max rel %.3e" % (err / np.maximum(np.abs(fd), 1e-12)).max())
    assert np.all(err <= 1e-4 * np.abs(fd) + 1e-6 * scale)


def test_slope_adjoint_matches_fp64_autograd():
    z = dem_for_gradients(150, 170, 31)
    t = targets(z, 32)
    g = gpu_grad(z, t)
    ref = torch_loss_grad(z.astype(np.float64), t, CS)
    ref[z == ND] = 0.0
    err = np.abs(g - ref)
    print("adjoint vs fp64 autograd: max abs %.3e (max |g| %.3e)" % (err.max(), np.abs(ref).max()))
    assert err.max() <= 1e-4 * np.abs(ref).max()


def test_gradient_descent_reduces_loss():
    import torch
    z0 = dem_for_gradients(256, 256, 41, nodata_discs=2)
    t = torch.from_numpy(targets(z0, 42).astype(np.float32)).cuda()
    dem = torch.from_numpy(z0).cuda()
    losses = []
    for _ in range(10):
        d = dem.clone().requires_grad_(True)
        s = autograd.slope(d, CS, ND)
        valid = s != ND
        loss = (((s - t) ** 2) * valid).sum()
        loss.backward()
        losses.append(loss.item())
        with torch.no_grad():
            step = d.grad * 200.0
            dem = torch.where(dem == ND, dem, dem - step)
    assert losses[-1] < 0.5 * losses[0], losses
