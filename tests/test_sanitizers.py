"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the host code
(CPU only):

- the GeoTIFF and point-shapefile readers (pargo_amd/gpro/geotiff.cpp,
  shapefile.cpp), which read untrusted files, under a seeded mutation fuzz
  of the reference's own data files (tests/cpp/fuzz_io.cpp);
- the CPU oracle (oracle/pargo_oracle.c) on small random inputs with NoData,
  NaN and infinities (tests/cpp/oracle_san.c);
- the gpro host layer's own checks (tests/cpp/gpro_host_test.cpp) with the
  codecs, the .nbr parser, the band cut and the layer classes built
  sanitized (libpargo.so supplies only the remaining C-ABI symbols).

Any out-of-bounds access, use after free, signed overflow, misaligned or
out-of-range conversion, or leak fails the run.  Only g++/gcc are used here:
no device code is built with a sanitizer.
"""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(ROOT, "tests", "golden", "data")
GPRO = os.path.join(ROOT, "pargo_amd", "gpro")
LIBDIR = os.path.join(ROOT, "pargo_amd", "_build")
SAN = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                                reason="needs the host compilers")


def _run(cmd, **kw):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, **kw)
    return p.returncode, p.stdout


@pytest.fixture(scope="module")
def san_bins(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("san"))
    jobs = {
        "fuzz_io": ["g++", "-std=c++17"] + SAN + [os.path.join(ROOT, "tests", "cpp", "fuzz_io.cpp"),
                                                  os.path.join(GPRO, "geotiff.cpp"), os.path.join(GPRO, "shapefile.cpp"),
                                                  os.path.join(ROOT, "pargo_amd", "csrc", "faults.cpp")],
        "oracle_san": ["gcc", "-std=c11", "-ffp-contract=off"] + SAN + [
            os.path.join(ROOT, "tests", "cpp", "oracle_san.c"), os.path.join(ROOT, "oracle", "pargo_oracle.c"), "-lm"],
        "gpro_host_san": ["g++", "-std=c++17", "-fopenmp", "-ffp-contract=off"] + SAN + [
            os.path.join(ROOT, "tests", "cpp", "gpro_host_test.cpp"), os.path.join(GPRO, "geotiff.cpp"),
            os.path.join(GPRO, "shapefile.cpp"), os.path.join(GPRO, "focalDevice.cpp"),
            os.path.join(GPRO, "deviceRaster.cpp"),
            os.path.join(GPRO, "application.cpp"), os.path.join(ROOT, "pargo_amd", "csrc", "dcmp.cpp"),
            os.path.join(ROOT, "pargo_amd", "csrc", "faults.cpp"),
            "-L" + LIBDIR, "-lpargo", "-Wl,-rpath," + LIBDIR],
    }

    def build(item):
        name, cmd = item
        rc, out = _run(cmd + ["-o", os.path.join(d, name)])
        return name, rc, out

    with ThreadPoolExecutor(max_workers=3) as ex:
        for name, rc, out in ex.map(build, jobs.items()):
            assert rc == 0, "sanitizer build of %s failed:\n%s" % (name, out[-4000:])
    return d


@pytest.mark.parametrize("seed", [1, 2])
def test_readers_survive_mutated_files(san_bins, seed, tmp_path):
    files = [os.path.join(DATA, f) for f in ("dem.tif", "mask_500.tif", "samples.shp")]
    rc, out = _run([os.path.join(san_bins, "fuzz_io"), str(tmp_path), "300", str(seed)] + files, env=ENV,
                   timeout=600)
    assert rc == 0, out[-6000:]
    assert "0 contract failures" in out


def test_oracle_under_sanitizers(san_bins):
    rc, out = _run([os.path.join(san_bins, "oracle_san"), "300", "7"], env=ENV, timeout=600)
    assert rc == 0, out[-6000:]


def test_gpro_host_layer_under_sanitizers(san_bins, tmp_path):
    rc, out = _run([os.path.join(san_bins, "gpro_host_san"), DATA, str(tmp_path)], env=ENV, timeout=600)
    assert rc == 0, out[-6000:]
    assert "gpro host checks passed" in out
