"""The Fortran drop-in modules build against the reference's own mod_global.

The modules of rumdeed_amd/fortran use only mod_global names that the
reference declares identically (src/mod_global.F90), so a RUMDEED program
keeps its own mod_global and swaps in mod_verlet, mod_pair and the emission
modules.  This compiles the reference's mod_global.F90 where it lies, then
the drop-in modules and their test driver against it, and links them with
librumdeed.so.  CPU only; skipped without flang or the reference tree."""
import os
import subprocess

import pytest

from conftest import ROOT

FORTRAN = os.path.join(ROOT, "rumdeed_amd", "fortran")
REF_GLOBAL = "/root/reference/src/mod_global.F90"


def test_dropin_modules_build_against_the_reference_mod_global(tmp_path):
    from rumdeed_amd.build import FORTRAN_MODULES, LIB, PKG, flang

    fc = flang()
    if fc is None or not os.path.exists(REF_GLOBAL):
        pytest.skip("needs flang and the reference tree")
    if not os.path.exists(LIB):
        pytest.skip("librumdeed.so not built")
    mods = [m for m in FORTRAN_MODULES if m != "mod_global.F90"]
    srcs = [REF_GLOBAL] + [os.path.join(FORTRAN, m) for m in mods]
    srcs.append(os.path.join(FORTRAN, "rd_fortran_tests.f90"))
    out = tmp_path / "rd_dropin_ref_global"
    res = subprocess.run([fc, "-O0", "-J", str(tmp_path), *srcs, "-L", PKG, "-lrumdeed",
                          "-o", str(out)], capture_output=True, text=True, cwd=tmp_path)
    assert res.returncode == 0, res.stdout + res.stderr
    assert out.exists()
