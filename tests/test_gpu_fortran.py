"""The ISO_C_BINDING layer (rumdeed_amd/fortran): the reference's test
scenarios run from Fortran through the drop-in routines."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "rumdeed_amd", "fortran", "rd_fortran_tests")


def test_fortran_dropin_layer():
    if not os.path.exists(BIN):
        pytest.skip("Fortran layer not built (flang unavailable at build time)")
    res = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "ALL PASSED" in res.stdout
