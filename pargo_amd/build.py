"""Builds libpargo.so (HIP kernels for gfx950 + host C ABI) and the CPU oracle.

Usage: python -m pargo_amd.build [--jobs N] [--no-oracle]

Outputs (git-ignored, shipped in-tree):
  pargo_amd/_build/libpargo.so       the product library behind include/pargo.h
  pargo_amd/_build/libpargo_ceiling.so  streaming kernels for bench.py's memory ceiling
  pargo_amd/_build/bin/{slope,flowdird8,idw}   drop-in command lines
  oracle/_build/liboracle.so         CPU oracle (test infrastructure only)
"""
import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "pargo_amd")
CSRC = os.path.join(PKG, "csrc")
APPS = os.path.join(PKG, "apps")
BUILD = os.path.join(PKG, "_build")
OBJ = os.path.join(BUILD, "obj")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

# -ffp-contract=off: the D8 path relies on each FP64 operation rounding where
# the source says; fused multiply-adds are written explicitly where wanted.
# packed FP32 instructions are disabled (known-bad on this hardware pool).
HIP_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-ffp-contract=off",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-ffp-contract=off"]

HIP_SOURCES = ["ctx.hip", "focal.hip", "idw.hip", "adjoint.hip", "sca.hip", "fill.hip", "compute.hip"]
CXX_SOURCES = ["synth.cpp", "dcmp.cpp", "io.cpp", "faults.cpp", "../gpro/geotiff.cpp", "../gpro/shapefile.cpp"]


def _run(cmd, cwd=None):
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + proc.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ... (exit %d)" % proc.returncode)
    return proc.stdout


def _compile(src):
    path = os.path.join(CSRC, src)
    out = os.path.join(OBJ, os.path.splitext(os.path.basename(src))[0] + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + ["-c", path, "-o", out]
    else:
        cmd = [CXX] + CXX_FLAGS + ["-c", path, "-o", out]
    _run(cmd)
    return out


def build_library(jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in HIP_SOURCES + CXX_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = os.path.join(BUILD, "libpargo.so.tmp")
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs + ["-lgomp"])
    os.replace(tmp, os.path.join(BUILD, "libpargo.so"))
    build_ceiling()
    return os.path.join(BUILD, "libpargo.so")


def build_ceiling():
    """libpargo_ceiling.so: the streaming kernels bench.py measures the
    memory system with (a measurement helper, not part of the product)."""
    os.makedirs(BUILD, exist_ok=True)
    tmp = os.path.join(BUILD, "libpargo_ceiling.so.tmp")
    _run([HIPCC] + HIP_FLAGS + ["-shared", os.path.join(CSRC, "ceiling.hip"), "-o", tmp])
    os.replace(tmp, os.path.join(BUILD, "libpargo_ceiling.so"))
    return os.path.join(BUILD, "libpargo_ceiling.so")


GPRO = os.path.join(PKG, "gpro")
# the GeoTIFF and shapefile codecs are part of libpargo.so (pargo_geotiff_*)
GPRO_SOURCES = ["application.cpp", "focalDevice.cpp", "hydroDevice.cpp", "deviceRaster.cpp"]
# command line -> (app directory, sources)
APP_SOURCES = {
    "slope": ("slope", ["slopeOperator.cpp", "pSlope.cpp"]),
    "flowdird8": ("flowdird8", ["d8Operator.cpp", "pD8.cpp"]),
    "idw": ("idw", ["idwOperator.cpp", "idwTransformation.cpp", "pIDW.cpp"]),
    "scad8": ("scad8", ["scaOperator.cpp", "pSCA.cpp"]),
    "pitremove": ("pitremove", ["pitRemoveOperator.cpp", "pPitRemove.cpp"]),
    # host-only checks of the gpro layer, run by tests/test_gpro_host.py
    "gpro_host_test": ("../../tests/cpp", ["gpro_host_test.cpp"]),
    # GPU checks of the device-resident layers, run by tests/test_gpu_gpro_chain.py
    "gpro_device_chain": ("../../tests/cpp", ["gpro_device_chain.cpp", "../../pargo_amd/apps/flowdird8/d8Operator.cpp",
                                              "../../pargo_amd/apps/scad8/scaOperator.cpp",
                                              "../../pargo_amd/apps/slope/slopeOperator.cpp"]),
}
APP_FLAGS = ["-O2", "-std=c++17", "-fopenmp", "-Wall", "-ffp-contract=off"]


def _compile_host(src, out):
    _run([CXX] + APP_FLAGS + ["-c", src, "-o", out])
    return out


def build_apps(jobs=8):
    """The gpro host layer and the slope / flowdird8 / idw command lines,
    linked against libpargo.so next to them (rpath $ORIGIN/..)."""
    aobj = os.path.join(OBJ, "apps")
    bindir = os.path.join(BUILD, "bin")
    os.makedirs(aobj, exist_ok=True)
    os.makedirs(bindir, exist_ok=True)
    jobs_list = [(os.path.join(GPRO, s), os.path.join(aobj, "gpro_" + s[:-4] + ".o")) for s in GPRO_SOURCES]
    for name, (d, srcs) in APP_SOURCES.items():
        jobs_list += [(os.path.normpath(os.path.join(APPS, d, s)),
                       os.path.join(aobj, name + "_" + os.path.basename(s)[:-4] + ".o")) for s in srcs]
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(lambda a: _compile_host(*a), jobs_list))
    gobjs = [os.path.join(aobj, "gpro_" + s[:-4] + ".o") for s in GPRO_SOURCES]
    outs = []
    for name, (d, srcs) in APP_SOURCES.items():
        objs = [os.path.join(aobj, name + "_" + os.path.basename(s)[:-4] + ".o") for s in srcs]
        exe = os.path.join(bindir, name)
        _run([CXX, "-o", exe + ".tmp"] + objs + gobjs +
             ["-L" + BUILD, "-lpargo", "-fopenmp", "-Wl,-rpath,$ORIGIN/.."])
        os.replace(exe + ".tmp", exe)
        outs.append(exe)
    return outs


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "_build", "liboracle.so")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--no-oracle", action="store_true")
    a = ap.parse_args(argv)
    print(build_library(a.jobs))
    for exe in build_apps(a.jobs):
        print(exe)
    if not a.no_oracle:
        print(build_oracle())


if __name__ == "__main__":
    main()
