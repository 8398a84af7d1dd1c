"""librumdeed.so loads and exports every entry point include/rumdeed.h
declares (no device calls)."""
import ctypes
import os
import re

from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "rumdeed.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rd_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    from rumdeed_amd import _lib

    assert sorted(_lib.exported_symbols()) == declared_symbols()


def test_library_exports_declared_symbols():
    from rumdeed_amd import _lib

    L = _lib.lib()
    for name in declared_symbols():
        assert hasattr(L, name), name
    assert L.rd_version().decode().startswith("rumdeed")


def test_errors_are_reported_without_a_device():
    from rumdeed_amd import _lib

    L = _lib.lib()
    h = ctypes.c_void_p()
    code = L.rd_ctx_create(0, 0, ctypes.byref(h))
    assert code == 1  # RD_ERR_ARG: max_particles outside range
    assert b"max_particles" in L.rd_last_error()
    assert L.rd_ctx_destroy(None) == 0
