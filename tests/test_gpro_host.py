"""Host-only checks of the C++ gpro layer (tests/cpp/gpro_host_test.cpp,
built by python -m pargo_amd.build) and of the row-band decomposition in
libpargo against the oracle.  No GPU is used."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po
from pargo_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "pargo_amd", "_build", "bin", "gpro_host_test")
DATA = os.path.join(ROOT, "tests", "golden", "data")


def test_gpro_host_checks(tmp_path):
    if not os.path.exists(BIN):
        pytest.fail("gpro_host_test not built: run python -m pargo_amd.build")
    p = subprocess.run([BIN, DATA, str(tmp_path)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stdout


def _bands(load, nodata, nsub):
    load = np.ascontiguousarray(load, dtype=np.float64)
    first = np.empty(nsub + 1, dtype=np.int32)
    rc = _lib.lib().pargo_load_row_bands(_lib._ptr(load), load.shape[0], load.shape[1], float(nodata), int(nsub),
                                         _lib._ptr(first))
    _lib.check_host(rc, "pargo_load_row_bands")
    return first


@pytest.mark.parametrize("seed", range(12))
def test_row_bands_match_oracle(seed):
    rng = np.random.default_rng(100 + seed)
    rows = int(rng.integers(1, 80))
    cols = int(rng.integers(1, 9))
    load = np.floor(rng.exponential(20.0, size=(rows, cols)))
    load[rng.random(load.shape) < 0.1] = -9999.0
    for nsub in sorted(n for n in {1, 2, 3, max(1, rows // 2), rows} if n <= rows):
        assert _bands(load, -9999.0, nsub).tolist() == po.val_row_dcmp(load, -9999.0, nsub).tolist()


def test_row_bands_reject_bad_counts():
    with pytest.raises(_lib.PargoError):
        _bands(np.ones((4, 2)), -9999.0, 5)
    with pytest.raises(_lib.PargoError):
        _bands(np.ones((4, 2)), -9999.0, 0)
