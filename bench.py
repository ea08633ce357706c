#!/usr/bin/env python3
"""Benchmark of the focal-operator hot path on one MI355X.

Workload (default): Horn slope + D8 flow direction (ESRI codes, dirType 256)
fused in one pass over a 16384 x 16384 float32 DEM with 30 m cells: seeded
fractal value-noise terrain, 2% plateaus, 1% NoData discs.  One step is one
pargo_slope_d8 call over the whole DEM, input already resident in HBM.

Prints one JSON line with Gcells/s, the HBM roofline fraction of the kernel
(9 algorithmic bytes per cell: 4 B DEM read, 4 B slope and 1 B code written)
and the CPU baseline (the FP64 restatement with one OpenMP row band per
thread, slope sweep + D8 sweep, on a band of rows of the same DEM).

Usage: python bench.py [--gpus 1] [--steps 20] [--warmup 5]
"""
import argparse
import ctypes
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = 9
NODATA = -9999.0
CELL = 30.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_threads():
    try:
        return max(1, int(os.environ.get("OMP_NUM_THREADS", "16")))
    except ValueError:
        return 16


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=16384)
    ap.add_argument("--cpu-rows", type=int, default=2048, help="rows of the DEM in the CPU baseline sample")
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from pargo_amd import _lib, focal, synth

    t0 = time.time()
    z = synth.workload_dem("slope_d8", a.size, a.size)
    nd_frac = float((z == NODATA).mean())
    log("generated %dx%d DEM in %.1fs (NoData %.4f)" % (a.size, a.size, time.time() - t0, nd_frac))

    dev = torch.device("cuda:0")
    dem = torch.from_numpy(z).to(dev)
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

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()

    # Timed region: K steps, events around every launch on the launch stream.
    stream = torch.cuda.current_stream(dev)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(a.steps)]
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    for s_ev, e_ev in evs:
        s_ev.record(stream)
        step()
        e_ev.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    kern_ms = [s.elapsed_time(e) for s, e in evs]
    avg_ms = sum(kern_ms) / len(kern_ms)
    ms_per_step = wall * 1000.0 / a.steps
    cells = a.size * a.size
    value = cells / (ms_per_step / 1000.0) / 1e9
    achieved = BYTES_PER_CELL * cells / (avg_ms / 1000.0) / 1e9
    log("kernel avg %.4f ms (min %.4f max %.4f), wall %.4f ms/step" % (avg_ms, min(kern_ms), max(kern_ms),
                                                                      ms_per_step))

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
        "dtype": "f32 in/out, f64 D8 decisions",
        "data": "synthetic (seeded fractal DEM, 2%% plateaus, %.2f%% NoData)" % (100 * nd_frac),
        "config": {"workload": "slope+D8 fused, Horn FD, dirType 256", "rows": a.size, "cols": a.size,
                   "cellsize_m": CELL, "dem_dtype": "float32", "outputs": "slope f32 + D8 code u8"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "focal_kernel<4,VEC,slope,d8,horn>", "kernel_ms": round(avg_ms, 5),
                     "algorithmic_bytes_per_cell": BYTES_PER_CELL},
    }

    # Parity spot check and CPU baseline on a band of rows of the same DEM.
    if not a.no_cpu_baseline:
        from oracle import pyoracle
        rows_s = min(a.cpu_rows, a.size)
        band = np.ascontiguousarray(z[:rows_s])
        nth = cpu_threads()
        times = []
        for _ in range(a.cpu_repeats):
            c0 = time.perf_counter()
            ref_s = pyoracle.slope(band, CELL, NODATA, "FD", nthreads=nth)
            ref_d = pyoracle.d8(band, NODATA, 256, nthreads=nth)
            times.append(time.perf_counter() - c0)
        tmed = statistics.median(times)
        ncell = rows_s * a.size
        out["cpu_baseline"] = {"value": round(ncell / tmed / 1e9, 4), "unit": "Gcells/s", "cores": nth,
                               "kind": "port",
                               "sample": "rows 0..%d of the same DEM (%d cells), slope sweep + D8 sweep, FP64, "
                                         "median of %d" % (rows_s - 1, ncell, a.cpu_repeats)}
        # compare interior rows of the band (the band's last row is a border there)
        gs = slope[: rows_s - 1].cpu().numpy()
        gc = code[: rows_s - 1].cpu().numpy().astype(np.int64)
        rs, rd = ref_s[: rows_s - 1], ref_d[: rows_s - 1]
        nd = rd == NODATA
        rc_ = np.where(rd == -1, 0, np.where(nd, 0, rd)).astype(np.int64)
        d8_ok = bool(np.array_equal(np.where(nd, 0, gc), rc_))
        v = rs != NODATA
        rel = np.abs(gs[v] - rs[v]) / np.maximum(np.abs(rs[v]), 1e-30)
        out["parity"] = {"d8_bitexact": d8_ok, "slope_max_rel": float(rel.max()),
                         "slope_rms_rel": float(np.sqrt(np.mean(rel ** 2))), "cells_checked": int(v.size)}
        out["speedup_vs_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
