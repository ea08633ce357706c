"""Device-resident layers of the gpro C++ layer (tests/cpp/gpro_device_chain.cpp,
built by python -m pargo_amd.build): DEM -> D8 -> SCA through the operator
API moves no byte between host and device between the operators, reuses the
DEM's device copy, and gives the host path's values bit for bit."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "pargo_amd", "_build", "bin", "gpro_device_chain")
DATA = os.path.join(ROOT, "tests", "golden", "data")


def test_device_chain():
    if not os.path.exists(BIN):
        pytest.fail("gpro_device_chain not built: run python -m pargo_amd.build")
    p = subprocess.run([BIN, os.path.join(DATA, "dem.tif"), os.path.join(DATA, "moore.nbr")],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(p.stdout)
    assert p.returncode == 0, p.stdout
    assert "gpro device chain checks passed" in p.stdout
