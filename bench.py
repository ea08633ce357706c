#!/usr/bin/env python3
"""Benchmark of the focal-operator hot path on one MI355X.

Headline workload: Horn slope + D8 flow direction (ESRI codes, dirType 256)
fused in one pass over a 16384 x 16384 float32 DEM with 30 m cells (seeded
fractal value-noise terrain, 2% plateaus, 1% NoData discs).  One step is one
pargo_slope_d8 call over the whole DEM, input already resident in HBM.

The JSON line carries Gcells/s, the HBM roofline of the fused kernel (9
algorithmic bytes per cell: 4 B DEM read, 4 B slope and 1 B code written)
next to the measured streaming ceiling of that traffic mix, and the CPU
baseline (the FP64 restatement, one OpenMP row band per thread, slope sweep
+ D8 sweep).  Parity of the timed output against the oracle covers every
cell: D8 codes and both NoData sets exactly, slope max/RMS relative error.
Unless --no-secondary is given it also measures, under "secondary":
  ceiling  streaming kernels with the fused pass's access pattern, no math
  adjoint  slope adjoint vs forward on an 8192^2 DEM (time ratio)
  idw      IDW of 20000 samples onto 4096^2 cells (K=12, p=2, 5 km)
  batch    512 uneven tiles (256..4096 cells a side) per-tile vs one graph
  sca      D8 contributing area (scad8) of the headline DEM's dirType-8 codes
  fill     depression filling (pitremove) of the headline DEM

Before the --warmup steps, the headline step runs for --settle-ms of wall
time so that the GPU clocks have left their idle state (the DEM is
generated on the host first, and the GPU idles meanwhile); the JSON line
states this under "clock_settle".  The adjoint, IDW, scad8 and pitremove legs
likewise run their kernels for 100-200 ms before timing them.

Usage: python bench.py [--gpus 1] [--steps 20] [--warmup 5]
"""
import argparse
import ctypes
import json
import math
import os
import re
import statistics
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
BYTES_PER_CELL = 9
NODATA = -9999.0
CELL = 30.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "focal_slope_d8_16384_hbm_traffic.txt")
CEILING_LIB = os.path.join(ROOT, "pargo_amd", "_build", "libpargo_ceiling.so")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """Host cores available to this process and how they were counted:
    the affinity set, capped by a cgroup CPU quota when one is set."""
    info = {"nproc": os.cpu_count()}
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    info["affinity_cpus"] = aff
    threads = aff
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = int(q) / int(per)
            info["cgroup_cpu_quota"] = round(quota, 2)
            threads = max(1, min(threads, int(math.floor(quota))))
    except (OSError, ValueError):
        pass
    try:
        txt = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True,
                             timeout=10).stdout
        for key, name in (("Socket(s)", "sockets"), ("Core(s) per socket", "cores_per_socket"),
                          ("Thread(s) per core", "threads_per_core"), ("Model name", "model")):
            m = re.search(r"^%s:\s*(.+)$" % re.escape(key), txt, re.M)
            if m:
                v = m.group(1).strip()
                info[name] = int(v) if v.isdigit() else v
        if "sockets" in info and "cores_per_socket" in info:
            info["physical_cores"] = info["sockets"] * info["cores_per_socket"]
    except (OSError, subprocess.SubprocessError):
        pass
    info["threads_used"] = threads
    return info


def settle(torch, fn, ms=200.0):
    """Runs fn for about ms of wall time so that the GPU clocks leave the
    state an idle spell (host-side data generation) left them in."""
    w0 = time.perf_counter()
    while (time.perf_counter() - w0) * 1000.0 < ms:
        for _ in range(10):
            fn()
        torch.cuda.synchronize()


def event_time(torch, fn, n, warm=2):
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(n):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / n


def measured_traffic():
    """HBM bytes per launch of the fused kernel from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), if present."""
    try:
        txt = open(TRAFFIC_FILE).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt)
    return float(m.group(1)) if m else None


def copy_probe():
    """A plain 1 GiB device copy (known traffic: 1 GiB read + 1 GiB write),
    the calibration run for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    ms = event_time(torch, lambda: b.copy_(a), 5)
    print(json.dumps({"copy_probe_bytes": 2 * 4 * n, "ms": ms, "GBs": 8 * n / ms / 1e6}), flush=True)


CEIL_VARIANTS = {"strip": (0, 9), "strip_halo": (1, 9), "flat": (2, 9), "strip_nt": (3, 9), "flat_nt": (6, 9),
                 "row1_halo": (7, 9), "stack8": (8, 9), "read": (4, 4), "write": (5, 5)}


def secondary_ceiling(torch, size):
    """Streaming kernels with the fused pass's traffic mix (4 B read, 4 B +
    1 B written per cell; 16-row strips, 4 columns per lane) and no math:
    what the memory system delivers for that pattern on this GPU."""
    L = ctypes.CDLL(CEILING_LIB)
    L.pargo_ceiling_launch.restype = ctypes.c_int
    L.pargo_ceiling_launch.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_int,
                                                                                ctypes.c_longlong, ctypes.c_int,
                                                                                ctypes.c_void_p]
    src = torch.rand((size, size), dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst8 = torch.empty((size, size), dtype=torch.uint8, device="cuda")
    sink = torch.zeros(1, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    cells = size * size
    res = {"config": "streaming kernels, %dx%d float32 in, float32 + uint8 out, no arithmetic" % (size, size)}
    best = 0.0
    for name, (v, bpc) in CEIL_VARIANTS.items():
        def go(v=v):
            rc = L.pargo_ceiling_launch(v, src.data_ptr(), dst.data_ptr(), dst8.data_ptr(), sink.data_ptr(), size,
                                        size, size, 16, stream)
            if rc:
                raise RuntimeError("pargo_ceiling_launch(%d) = %d" % (v, rc))
        ms = event_time(torch, go, 20, warm=5)
        gbs = bpc * cells / ms / 1e6
        res[name + "_GBs"] = round(gbs, 1)
        if bpc == 9:
            best = max(best, gbs)
    res["ceiling_9B_GBs"] = round(best, 1)
    res["ceiling_frac_of_8TBs"] = round(best / HBM_PEAK_GBS, 4)
    del src, dst, dst8, sink
    torch.cuda.empty_cache()
    return res


def secondary_adjoint(torch, _lib, focal, synth):
    N = 8192
    z = synth.dem(synth.BASE_SEED + 3, N, N)
    synth.add_nodata_discs(synth.BASE_SEED + 3, z, fraction=0.01)
    dem = torch.from_numpy(z).cuda()
    del z
    s = torch.empty_like(dem)
    g = torch.rand_like(dem)
    gd = torch.empty_like(dem)
    ctx = _lib.default_context(0)
    ctx.bind_stream()
    L = _lib.lib()
    grid = focal.grid_of(dem, CELL, NODATA)
    def fwd_step():
        ctx.check(L.pargo_slope(ctx.h, ctypes.byref(grid), _lib._ptr(dem), 0, 0, _lib._ptr(s), None))

    def adj_step():
        ctx.check(L.pargo_slope_adjoint(ctx.h, ctypes.byref(grid), _lib._ptr(dem), _lib._ptr(g), _lib._ptr(gd)))

    # both at settled clocks (the DEM's generation leaves the GPU idle); the
    # median of three timings of 20 launches each
    settle(torch, fwd_step)
    fwd = sorted(event_time(torch, fwd_step, 20) for _ in range(3))[1]
    settle(torch, adj_step)
    adj = sorted(event_time(torch, adj_step, 20) for _ in range(3))[1]
    out = {"config": "Horn slope and its adjoint on an 8192^2 float32 DEM (1% NoData)",
           "timing": "each after 200 ms of its own launches, median of 3 x 20 launches",
           "fwd_ms": round(fwd, 5), "adjoint_ms": round(adj, 5), "adjoint_fwd_ratio": round(adj / fwd, 3),
           "fwd_GBs": round(8 * N * N / fwd / 1e6, 1), "adjoint_GBs": round(12 * N * N / adj / 1e6, 1),
           "adjoint_bytes_per_cell": 12}
    del dem, s, g, gd
    torch.cuda.empty_cache()
    return out


def secondary_methods(torch, _lib, focal, z, cpu):
    """Every -mtd slope method on the headline DEM: fused slope+D8
    (dirType 256) and slope alone, each timed over 10 launches.  With the CPU
    leg, the oracle's slope of a 256-row band is timed per method (one
    thread) and the device slope checked against it there (relative 1e-5,
    NoData set equal)."""
    dem = torch.from_numpy(z).cuda()
    rows, cols = z.shape
    s = torch.empty_like(dem)
    c = torch.empty(dem.shape, dtype=torch.uint8, device="cuda")
    ctx = _lib.default_context(0)
    ctx.bind_stream()
    L = _lib.lib()
    grid = focal.grid_of(dem, CELL, NODATA)
    r0 = rows // 2
    res = {"config": "slope per -mtd method on the %dx%d headline DEM, tan units; fused = slope + D8 dirType 256"
                     % (rows, cols)}
    if cpu is not None:
        from oracle import pyoracle
        import numpy as np
        band = z[r0 - 1:r0 + 257]
        res["cpu_sample"] = "rows %d..%d (a 258-row band with its halo rows), one thread" % (r0, r0 + 255)
    names = ("FD", "FFD", "MD", "SD", "SFD", "TFD", "TFDW")
    for m, name in enumerate(names):
        fused = event_time(torch, lambda: ctx.check(L.pargo_slope_d8(ctx.h, ctypes.byref(grid), _lib._ptr(dem), m, 0,
                                                                     256, _lib._ptr(s), _lib._ptr(c), None)), 10)
        alone = event_time(torch, lambda: ctx.check(L.pargo_slope(ctx.h, ctypes.byref(grid), _lib._ptr(dem), m, 0,
                                                                  _lib._ptr(s), None)), 10)
        r = {"fused_ms": round(fused, 4), "fused_Gcells_s": round(rows * cols / fused / 1e6, 1),
             "slope_ms": round(alone, 4), "slope_Gcells_s": round(rows * cols / alone / 1e6, 1),
             "slope_GBs": round(8.0 * rows * cols / alone / 1e6, 1)}
        if cpu is not None:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ref = pyoracle.slope(band, CELL, NODATA, name)[1:-1, 1:-1]
            dt = time.perf_counter() - t0
            got = s[r0:r0 + 256, 1:-1].cpu().numpy()
            nd = ref == NODATA
            r["cpu_Gcells_s"] = round(256 * (cols - 2) / dt / 1e9, 4)
            r["band_within_1e-5"] = bool(np.array_equal(got == np.float32(NODATA), nd)) and bool(
                np.all(np.abs(got[~nd] - ref[~nd]) <= 1e-5 * np.abs(ref[~nd]) + 1e-12))
        res[name] = r
    del dem, s, c
    torch.cuda.empty_cache()
    return res


def secondary_idw(torch, synth, IDWOperator, cpu):
    xs, ys, vs = synth.samples(synth.BASE_SEED + 2, 20000)
    op = IDWOperator(CELL, 12, 2, 5000.0, 5000.0).read_samples(xs, ys, vs)
    out, cand = op.run(count_candidates=True)
    torch.cuda.synchronize()
    c = int(cand.item())
    # timings at settled clocks: 200 ms of launches first, then the median
    # of three timings of 5 launches
    settle(torch, lambda: op.run(out=out))
    ms = sorted(event_time(torch, lambda: op.run(out=out), 5, warm=1) for _ in range(3))[1]
    cells = op.rows * op.cols
    flop = 5.0 * c + 49.0 * cells
    res = {"config": "IDW of 20000 samples onto 4096^2 cells, K=12, p=2, searchRange 5000, blockSize 5000",
           "ms": round(ms, 4), "Gcells_s": round(cells / ms / 1e6, 3), "candidates_per_cell": round(c / cells, 2),
           "roofline": {"bound": "fp64 valu", "achieved": round(flop / ms / 1e9, 3), "peak": FP64_PEAK_TFLOPS,
                        "unit": "TFLOP/s", "frac": round(flop / ms / 1e9 / FP64_PEAK_TFLOPS, 4),
                        "flop_per_launch": flop}}
    # tile launch order from the reference's load models (heaviest first):
    # the values must not change, only the time
    ref_out = out.clone()
    orders = {}
    for source in ("estimate", "measured"):
        op.order_tiles(source)
        o2 = op.run()
        torch.cuda.synchronize()
        same = bool(torch.equal(o2, ref_out))
        oms = sorted(event_time(torch, lambda: op.run(out=o2), 5, warm=1) for _ in range(3))[1]
        orders[source] = {"ms": round(oms, 4), "values_identical": same}
    op.order_tiles("rows")
    res["tile_order"] = {"rows_ms": res["ms"], **orders}
    # the IDW adjoint (dL/d sample values) for an L2 loss against a seeded
    # target grid: dL/d(out) = 2 (out - target)
    gen = torch.Generator(device="cuda").manual_seed(synth.BASE_SEED + 3)
    target = out + torch.randn(out.shape, device="cuda", generator=gen)
    grad = (2.0 * (out - target)).contiguous()
    settle(torch, lambda: op.adjoint(grad), 100.0)
    ams = sorted(event_time(torch, lambda: op.adjoint(grad), 3, warm=1) for _ in range(3))[1]
    res["adjoint_ms"] = round(ams, 4)
    res["adjoint_fwd_ratio"] = round(ams / ms, 3)
    if cpu is not None:
        from oracle import pyoracle
        import numpy as np
        # bounded sample: every 16th row of the grid (a rows x 1 mask layer
        # with 1 on those rows: the reference skips masked cells before
        # searching), compared with the GPU's rows
        stride = 16
        mask = np.zeros((op.rows, 1), dtype=np.int32)
        mask[::stride, 0] = 1
        nth = cpu["threads_used"]
        t0 = time.perf_counter()
        ref = pyoracle.idw(xs, ys, vs, CELL, k=12, power=2, buffer=5000.0, block_size=5000.0, mask=mask,
                           mask_nodata=-1, nthreads=nth)
        dt = time.perf_counter() - t0
        got = out.cpu().numpy().astype(np.float64)[::stride]
        r = ref[::stride]
        rel = np.abs(got - r) / np.maximum(np.abs(r), 1e-30)
        res["cpu_baseline"] = {"value": round(r.size / dt / 1e9, 6), "unit": "Gcells/s", "cores": nth,
                               "kind": "port", "sample": "every %dth row of the grid (%d cells; other rows masked "
                                                         "out)" % (stride, r.size)}
        res["parity_on_sample"] = {"max_rel": float(rel.max()), "rms_rel": float(np.sqrt(np.mean(rel ** 2))),
                                   "within_1e-5": bool(np.all(np.abs(got - r) <= 1e-5 * np.abs(r) + 1e-9)),
                                   "cells_checked": int(r.size)}
    op.free()
    return res


def secondary_sca(torch, synth, focal, hydro, z, cpu):
    """D8 contributing area (scad8) from the dirType-8 codes of the
    headline DEM, the reference's single sweep and the converging variant."""
    rows, cols = z.shape
    dem = torch.from_numpy(z).cuda()
    code, mask = focal.d8(dem, CELL, NODATA, dirtype=8, want_mask=True)
    del dem
    res = {"config": "D8 contributing area of the 16384^2 headline DEM (dirType-8 codes + NoData mask in, "
                     "float64 areas out)"}
    cells = rows * cols
    for mode in ("reference", "converge"):
        settle(torch, lambda: hydro.sca_d8(code, NODATA, mask, mode))
        ms = sorted(event_time(torch, lambda: hydro.sca_d8(code, NODATA, mask, mode), 5, warm=1)
                    for _ in range(3))[1]
        res[mode + "_ms"] = round(ms, 3)
        res[mode + "_Gcells_s"] = round(cells / ms / 1e6, 2)
    # 1 B code + 1/8 B mask in, 8 B area out per cell
    gbs = 9.125 * cells / res["reference_ms"] / 1e6
    res["roofline"] = {"bound": "hbm (random 8-byte hops: one per chain link)", "achieved": round(gbs, 1),
                       "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                       "algorithmic_bytes_per_cell": 9.125}
    if cpu is not None:
        from oracle import pyoracle
        import numpy as np
        band = 768  # rows of the D8 layer as their own grid
        d8 = focal.codes_to_reference(code[:band], focal.unpack_mask(mask[:band], cols), 8, NODATA).cpu().numpy()
        t0 = time.perf_counter()
        ref, _ = pyoracle.sca_d8(d8, NODATA, False)
        dt = time.perf_counter() - t0
        flow = hydro.d8_classify(torch.from_numpy(d8).cuda(), NODATA)
        got = hydro.sca_d8(flow, NODATA).cpu().numpy()
        res["cpu_baseline"] = {"value": round(band * cols / dt / 1e9, 5), "unit": "Gcells/s", "cores": 1,
                               "kind": "port", "sample": "rows 0..%d of the D8 layer as a grid of its own, "
                                                         "reference mode (the sweep order is sequential)" % (band - 1)}
        res["parity_bitexact_on_sample"] = bool(np.array_equal(got, ref))
    del code, mask
    torch.cuda.empty_cache()
    return res


def secondary_fill(torch, synth, hydro, z, cpu):
    """Depression filling (pitremove) of the headline DEM."""
    rows, cols = z.shape
    dem = torch.from_numpy(z).cuda()
    settle(torch, lambda: hydro.pit_remove(dem, NODATA))
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        w, passes = hydro.pit_remove(dem, NODATA)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    ms = min(times) * 1e3
    cells = rows * cols
    filled = int((w > dem.double()).sum().item())
    res = {"config": "pitremove of the 16384^2 headline DEM (float32 in, float64 out)", "ms": round(ms, 3),
           "Gcells_s": round(cells / ms / 1e6, 2), "device_passes": passes, "cells_raised": filled}
    gbs = 12.0 * cells / ms / 1e6  # 4 B DEM in, 8 B filled value out
    res["roofline"] = {"bound": "hbm (repeated tile passes until nothing changes)", "achieved": round(gbs, 1),
                       "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                       "algorithmic_bytes_per_cell": 12}
    del w, dem
    if cpu is not None:
        from oracle import pyoracle
        import numpy as np
        side = 384  # a window of the DEM as its own grid
        win = np.ascontiguousarray(z[:side, :side])
        t0 = time.perf_counter()
        ref, sweeps, ok = pyoracle.pit_remove(win, NODATA)
        dt = time.perf_counter() - t0
        got, _ = hydro.pit_remove(torch.from_numpy(win).cuda(), NODATA)
        res["cpu_baseline"] = {"value": round(side * side / dt / 1e9, 6), "unit": "Gcells/s", "cores": 1,
                               "kind": "port", "sample": "the %dx%d top-left window as a grid of its own "
                                                         "(%d Gauss-Seidel sweeps)" % (side, side, sweeps)}
        res["parity_bitexact_on_sample"] = bool(ok and np.array_equal(got.cpu().numpy(), ref))
    torch.cuda.empty_cache()
    return res


def secondary_batch(torch, batch, cpu):
    import numpy as np
    shapes = batch.tile_shapes(512, batch.synth.BASE_SEED + 1000)
    host = [batch.tile_dem(t, r, c) for t, (r, c) in enumerate(shapes)]
    dems = [torch.from_numpy(z).cuda() for z in host]
    cells = sum(r * c for r, c in shapes)
    b = batch.Batch(dems)
    per = event_time(torch, lambda: b.run(batch.PER_TILE), 3, warm=1)
    ref_s = [x.clone() for x in b.slopes]
    ref_c = [x.clone() for x in b.codes]
    cpt, rows = b.autotune()
    # graph replays for a fixed wall time first (the host-side tile setup
    # leaves the GPU idle and its clocks low), then the median of 3 timings
    w0 = time.perf_counter()
    while time.perf_counter() - w0 < 0.3:
        b.run(batch.GRAPH)
        torch.cuda.synchronize()
    gr = sorted(event_time(torch, lambda: b.run(batch.GRAPH), 5, warm=1) for _ in range(3))[1]
    # the graph's outputs against the per-tile launches (the autotuned shape
    # may differ from the per-tile one): D8 bit for bit, slope to the bit as
    # well (same arithmetic per cell)
    for x in b.slopes:
        x.fill_(float("nan"))
    for x in b.codes:
        x.fill_(77)
    b.run(batch.GRAPH)
    torch.cuda.synchronize()
    d8_same = all(torch.equal(b.codes[i], ref_c[i]) for i in range(len(dems)))
    slope_same = all(torch.equal(b.slopes[i], ref_s[i]) for i in range(len(dems)))
    # queue order from the load models (the default is the estimate, cells)
    orders = {"estimate_ms": round(gr, 3)}
    for source in ("rows", "measured"):
        b.order(source)
        ms = sorted(event_time(torch, lambda: b.run(batch.GRAPH), 5, warm=1) for _ in range(3))[1]
        orders[source + "_ms"] = round(ms, 3)
        orders[source + "_values_identical"] = all(torch.equal(b.codes[i], ref_c[i]) and
                                                   torch.equal(b.slopes[i], ref_s[i]) for i in range(len(dems)))
    b.order("estimate")
    res = {"config": "512 tiles, side exp(U[ln 256, ln 4096]), aspect U[0.5, 2], 1-5% NoData, slope+D8",
           "tiles": 512, "Gcells": round(cells / 1e9, 4), "total_bytes_in": 4 * cells,
           "total_bytes_algorithmic": 9 * cells, "per_tile_ms": round(per, 3), "graph_ms": round(gr, 3),
           "speedup": round(per / gr, 2), "graph_Gcells_s": round(cells / gr / 1e6, 1),
           "graph_GBs": round(9 * cells / gr / 1e6, 1), "autotuned": {"cols_per_lane": cpt, "rows_per_strip": rows},
           "graph_equals_per_tile": {"d8": d8_same, "slope": slope_same}, "queue_order": orders}
    if cpu is not None:
        from oracle import pyoracle
        # D8 and slope of every 32nd tile against the oracle
        ok_d8 = ok_s = True
        checked = 0
        for t in range(0, len(dems), 32):
            z = host[t]
            rd = pyoracle.d8(z, NODATA, 256, nthreads=cpu["threads_used"])
            rs = pyoracle.slope(z, CELL, NODATA, nthreads=cpu["threads_used"])
            nd = rd == NODATA
            code = np.where(rd == -1, 0, np.where(nd, 0, rd)).astype(np.int64)
            gc = b.codes[t].cpu().numpy().astype(np.int64)
            gs = b.slopes[t].cpu().numpy().astype(np.float64)
            ok_d8 &= bool(np.array_equal(np.where(nd, 0, gc), code)) and bool(np.all(gc[nd] == 0))
            v = rs != NODATA
            ok_s &= bool(np.array_equal(gs == NODATA, ~v))
            ok_s &= bool(np.all(np.abs(gs[v] - rs[v]) <= 1e-5 * np.abs(rs[v]) + 1e-12))
            checked += z.size
        res["oracle_on_sample"] = {"tiles": "every 32nd (16 tiles)", "cells": checked, "d8_bitexact": ok_d8,
                                   "slope_within_1e-5_and_nodata_equal": ok_s}
    del b, dems
    torch.cuda.empty_cache()
    return res


def headline_parity(torch, focal, pyoracle, np, dem, slope, code, cpu):
    """The timed call's outputs against the oracle on every cell: D8 codes
    bit for bit on valid cells and 0 on NoData ones, the D8 NoData set (the
    mask of one extra call) and the slope NoData set equal to the oracle's,
    slope max/RMS relative error on valid cells."""
    z = dem.cpu().numpy()
    nth = cpu["threads_used"]
    ref_s = pyoracle.slope(z, CELL, NODATA, "FD", nthreads=nth)
    ref_d = pyoracle.d8(z, NODATA, 256, nthreads=nth)
    _, _, mask = focal.slope_d8(dem, CELL, NODATA, want_slope=False, want_mask=True)
    gmask = focal.unpack_mask(mask, z.shape[1]).cpu().numpy()
    gs = slope.cpu().numpy()
    gc = code.cpu().numpy()
    nd8 = ref_d == NODATA
    nds = ref_s == NODATA
    d8_nodata_equal = bool(np.array_equal(gmask, nd8))
    rc_ = np.where(ref_d == -1, 0, np.where(nd8, 0, ref_d)).astype(np.uint8)
    d8_ok = bool(np.array_equal(gc, rc_))  # NoData cells must hold 0 as well
    slope_nodata_equal = bool(np.array_equal(gs == np.float32(NODATA), nds))
    v = ~nds
    rel = np.abs(gs[v].astype(np.float64) - ref_s[v]) / np.maximum(np.abs(ref_s[v]), 1e-30)
    return {"d8_bitexact": d8_ok and d8_nodata_equal, "d8_nodata_set_equal": d8_nodata_equal,
            "slope_nodata_set_equal": slope_nodata_equal, "slope_max_rel": float(rel.max()),
            "slope_rms_rel": float(np.sqrt(np.mean(rel ** 2))),
            "slope_within_1e-5": bool(np.all(np.abs(gs[v] - ref_s[v]) <= 1e-5 * np.abs(ref_s[v]) + 1e-12)),
            "cells": int(z.size), "valid_cells_checked": int(v.sum()), "nodata_cells": int(nds.sum())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=60)
    ap.add_argument("--settle-ms", type=float, default=400.0,
                    help="wall time of headline steps run before the warm-up steps (GPU clocks leave idle)")
    ap.add_argument("--size", type=int, default=16384)
    ap.add_argument("--cpu-repeats", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--only", choices=("ceiling", "adjoint", "idw", "batch", "sca", "fill", "methods"),
                    default=None,
                    help="run the headline and this one secondary workload only (for profiling)")
    ap.add_argument("--copy-probe", action="store_true", help="only run the 1 GiB copy used to calibrate counters")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path; "
              "a larger DEM would be split into row bands with a one-row halo, one band per GPU",
              file=sys.stderr)
        sys.exit(2)
    if a.copy_probe:
        copy_probe()
        return

    import numpy as np
    import torch
    from pargo_amd import _lib, batch, focal, hydro, synth
    from pargo_amd.idw import IDWOperator

    t0 = time.time()
    z = synth.workload_dem("slope_d8", a.size, a.size)
    nd_frac = float((z == NODATA).mean())
    log("generated %dx%d DEM in %.1fs (NoData %.4f)" % (a.size, a.size, time.time() - t0, nd_frac))

    dev = torch.device("cuda:0")
    dem = torch.from_numpy(z).to(dev)
    del z
    slope = torch.empty_like(dem)
    code = torch.empty(dem.shape, dtype=torch.uint8, device=dev)
    ctx = _lib.default_context(0)
    ctx.bind_stream()
    L = _lib.lib()
    g = focal.grid_of(dem, CELL, NODATA)
    gp = ctypes.byref(g)
    pd, ps, pc = _lib._ptr(dem), _lib._ptr(slope), _lib._ptr(code)

    def step():
        rc = L.pargo_slope_d8(ctx.h, gp, pd, 0, 0, 256, ps, pc, None)
        if rc:
            ctx.check(rc)

    # clock settling: headline steps for a fixed wall time, in bursts of 20
    settle_n = 0
    w0 = time.perf_counter()
    while (time.perf_counter() - w0) * 1000.0 < a.settle_ms:
        for _ in range(20):
            step()
        settle_n += 20
        torch.cuda.synchronize()
    settle_ms = (time.perf_counter() - w0) * 1000.0
    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()

    # Timed region: K back-to-back steps bracketed by synchronisation, with
    # HIP events at its two ends on the stream the kernels run on (torch's
    # current stream, which the context is bound to).  An event between
    # every step would add a timestamp packet and a cache flush to each one.
    stream = torch.cuda.current_stream(dev)
    s_ev, e_ev = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    s_ev.record(stream)
    for _ in range(a.steps):
        step()
    e_ev.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    avg_ms = s_ev.elapsed_time(e_ev) / a.steps
    ms_per_step = wall * 1000.0 / a.steps
    cells = a.size * a.size
    value = cells / (ms_per_step / 1000.0) / 1e9
    achieved = BYTES_PER_CELL * cells / (avg_ms / 1000.0) / 1e9
    log("step avg %.4f ms (events), wall %.4f ms/step" % (avg_ms, ms_per_step))

    out = {
        "metric": "Gcells/s slope+D8 on 16384^2 DEM, % HBM roofline; adjoint/fwd time ratio",
        "value": round(value, 3),
        "unit": "Gcells/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded fractal DEM, 2%% plateaus, %.2f%% NoData)" % (100 * nd_frac),
        "config": {"workload": "slope+D8 fused (Horn FD, tan units; D8 dirType 256) on a 16384x16384 float32 DEM",
                   "rows": a.size, "cols": a.size, "cellsize_m": CELL,
                   "io": "f32 DEM in, f32 slope + u8 D8 code out",
                   "arithmetic": "f32 storage; Horn numerators in f64, square root in f32; D8 by an f32 filter "
                                 "with a proven bound, undecided cells on the reference's f64 drops (codes "
                                 "identical to f64)"},
        "clock_settle": {"ms": round(settle_ms, 1), "launches": settle_n,
                         "note": "headline steps run before the warm-up steps so that the clocks leave idle"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(),
                     "kernel": "focal_kernel_fused<4,true> + focal_fixup_kernel (empty queue)",
                     "step_ms": round(avg_ms, 5), "algorithmic_bytes_per_cell": BYTES_PER_CELL},
    }

    cpu = None if a.no_cpu_baseline else host_cpu()
    if cpu is not None:
        from oracle import pyoracle
        zh = dem.cpu().numpy()
        nth = cpu["threads_used"]
        times = []
        for _ in range(a.cpu_repeats):
            c0 = time.perf_counter()
            pyoracle.slope(zh, CELL, NODATA, "FD", nthreads=nth)
            pyoracle.d8(zh, NODATA, 256, nthreads=nth)
            times.append(time.perf_counter() - c0)
        tmed = statistics.median(times)
        one = np.ascontiguousarray(zh[:512])
        c0 = time.perf_counter()
        pyoracle.slope(one, CELL, NODATA, "FD", nthreads=1)
        pyoracle.d8(one, NODATA, 256, nthreads=1)
        t1 = time.perf_counter() - c0
        del zh, one
        out["cpu_baseline"] = {"value": round(cells / tmed / 1e9, 4), "unit": "Gcells/s", "cores": nth,
                               "kind": "port",
                               "sample": "the whole %dx%d DEM, slope sweep + D8 sweep in FP64, one row band per "
                                         "OpenMP thread, median of %d" % (a.size, a.size, a.cpu_repeats),
                               "one_thread_Gcells_s": round(512 * a.size / t1 / 1e9, 4), "host": cpu}
        out["parity"] = headline_parity(torch, focal, pyoracle, np, dem, slope, code, cpu)
        out["speedup_vs_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    del dem, slope, code
    torch.cuda.empty_cache()

    if not a.no_secondary:
        sec = {}
        t0 = time.time()
        want = (lambda name: a.only in (None, name))
        if want("ceiling"):
            sec["ceiling"] = secondary_ceiling(torch, a.size)
            out["roofline"]["measured_ceiling_GBs"] = sec["ceiling"]["ceiling_9B_GBs"]
            out["roofline"]["frac_of_measured_ceiling"] = round(achieved / sec["ceiling"]["ceiling_9B_GBs"], 4)
            log("ceiling done (%.1fs)" % (time.time() - t0))
        if want("adjoint"):
            sec["adjoint"] = secondary_adjoint(torch, _lib, focal, synth)
            log("adjoint done (%.1fs)" % (time.time() - t0))
        if want("idw"):
            sec["idw"] = secondary_idw(torch, synth, IDWOperator, cpu)
            log("idw done (%.1fs)" % (time.time() - t0))
        if want("batch"):
            sec["batch"] = secondary_batch(torch, batch, cpu)
            log("batch done (%.1fs)" % (time.time() - t0))
        if want("sca") or want("fill") or want("methods"):
            zh = synth.workload_dem("slope_d8")
            if want("methods"):
                sec["methods"] = secondary_methods(torch, _lib, focal, zh, cpu)
                log("methods done (%.1fs)" % (time.time() - t0))
            if want("sca"):
                sec["sca"] = secondary_sca(torch, synth, focal, hydro, zh, cpu)
                log("sca done (%.1fs)" % (time.time() - t0))
            if want("fill"):
                sec["fill"] = secondary_fill(torch, synth, hydro, zh, cpu)
                log("fill done (%.1fs)" % (time.time() - t0))
            del zh
        out["secondary"] = sec
        if "adjoint" in sec:
            out["adjoint_fwd_ratio"] = sec["adjoint"]["adjoint_fwd_ratio"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
