"""Host-side argument checks of the Python mirror that must refuse a
tensor layout before any device work (CPU only)."""
import pytest
import torch

from pargo_amd import _lib, batch, focal, hydro


def test_expanded_dem_is_refused_before_launch():
    dem = torch.zeros(40).expand(30, 40)  # stride (0, 1): every row aliases one
    assert dem.stride() == (0, 1)
    with pytest.raises(ValueError, match="overlap"):
        _lib.row_pitch(dem, "dem")
    with pytest.raises(ValueError, match="overlap"):
        focal._check_dem(dem)
    with pytest.raises(ValueError, match="overlap"):
        focal.grid_of(dem, 30.0, -9999.0)
    with pytest.raises(ValueError, match="overlap"):
        focal._alloc_like(dem, torch.float32)
    with pytest.raises(ValueError, match="overlap"):
        hydro._check(dem, torch.float32, "dem")


def test_rows_must_be_contiguous():
    dem = torch.zeros(40, 30).t()  # stride (1, 30)
    with pytest.raises(ValueError, match="contiguous"):
        _lib.row_pitch(dem, "dem")


def test_pitched_view_keeps_its_pitch():
    big = torch.zeros(20, 64)
    view = big[:, 3:50]
    assert _lib.row_pitch(view) == 64
    assert focal.grid_of(view, 30.0, -9999.0).pitch == 64
    out = focal._alloc_like(view, torch.uint8)
    assert out.shape == view.shape and out.stride(0) == 64


def test_single_row_pitch_is_cols():
    row = torch.zeros(25).expand(1, 25)
    assert _lib.row_pitch(row) == 25


def test_wrong_dtype_and_cpu_tensor_are_refused():
    with pytest.raises(TypeError, match="float32"):
        focal._check_dem(torch.zeros(5, 5, dtype=torch.float64))
    with pytest.raises(TypeError, match="HIP device"):
        focal._check_dem(torch.zeros(5, 5))


def test_batch_refuses_expanded_tile():
    with pytest.raises(ValueError, match="overlap"):
        batch.Batch([torch.zeros(16).expand(8, 16)])
