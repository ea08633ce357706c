"""The C-ABI library loads and exports every entry point include/pargo.h
declares (no GPU calls), and the host-only parts behave."""
import hashlib
import os
import re

import numpy as np

from pargo_amd import _lib, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "pargo.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pargo_\w+)\s*\(", txt)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for n in ("pargo_slope_d8", "pargo_slope", "pargo_d8", "pargo_idw_run", "pargo_slope_adjoint",
              "pargo_idw_adjoint", "pargo_create", "pargo_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    missing = [n for n in declared_functions() if getattr(L, n, None) is None]
    assert not missing, "libpargo.so lacks %s" % missing


def test_binding_covers_header():
    assert set(declared_functions()) <= set(_lib.exported_symbols())


def test_abi_version():
    assert _lib.lib().pargo_abi_version() == 1


def test_null_context_errors_without_gpu():
    L = _lib.lib()
    g = _lib.PargoGrid(rows=10, cols=10, pitch=0, cellsize=30.0, nodata=-9999.0)
    import ctypes
    assert L.pargo_slope_d8(None, ctypes.byref(g), None, 0, 0, 256, None, None, None) == _lib.PARGO_E_INVALID
    assert L.pargo_sync(None) == _lib.PARGO_E_INVALID


def test_synthetic_dem_is_deterministic():
    a = synth.dem(123, 64, 80)
    b = synth.dem(123, 64, 80, nthreads=3)
    c = synth.dem(124, 64, 80)
    assert a.dtype == np.float32 and a.shape == (64, 80)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)
    assert a.min() == 0.0 and a.max() == 2000.0
    h = hashlib.sha256(a.tobytes()).hexdigest()
    assert h == hashlib.sha256(synth.dem(123, 64, 80).tobytes()).hexdigest()


def test_plateaus_and_nodata_fractions():
    z = synth.dem(5, 300, 300)
    synth.add_plateaus(5, z, 0.02)
    n = synth.add_nodata_discs(5, z, fraction=0.01)
    assert n == int((z == -9999.0).sum())
    assert 0.01 <= n / z.size < 0.05


def test_samples_extent_and_types():
    xs, ys, vs = synth.samples(synth.BASE_SEED + 2, 2000)
    assert xs[0] == 15.0 and ys[0] == 15.0 and xs[1] == 122865.0 and ys[1] == 122865.0
    assert np.all(xs.astype(np.float32) == xs) and np.all(vs.astype(np.float32) == vs)
    assert np.all((xs >= 15.0) & (xs <= 122865.0) & (ys >= 15.0) & (ys <= 122865.0))
