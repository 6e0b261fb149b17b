"""The C-ABI library loads and exports every entry point include/mmx.h declares
(no GPU needed: nothing here launches a kernel)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mmx.h")
LIB = os.path.join(ROOT, "micromagnetic.jl_amd", "libmmx.so")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*|void)\s*\*?\s*(mmx_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("mmx_create", "mmx_destroy", "mmx_effective_field", "mmx_llg_rhs", "mmx_advance_step",
                 "mmx_run_until", "mmx_relax", "mmx_add_demag", "mmx_add_exch", "mmx_add_dmi"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libmmx.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libmmx.so not built")
def test_binding_covers_header():
    import mmx._lib as L

    names = set(declared()) - {"mmx_last_error"}
    assert names == set(L.SIGNATURES), names ^ set(L.SIGNATURES)
    L.load()  # loads without a GPU


@pytest.mark.skipif(not os.path.exists(LIB), reason="libmmx.so not built")
def test_invalid_arguments_fail_loudly():
    import mmx._lib as L

    lib = L.load()
    d = L.MeshDesc(0, 1, 1, 0, 1e-9, 1e-9, 1e-9, 64, 0)
    h = ctypes.c_void_p()
    assert lib.mmx_create(ctypes.byref(h), ctypes.byref(d)) == L.EINVAL
    assert b"mesh sizes" in lib.mmx_last_error()
    d = L.MeshDesc(4, 4, 1, 0, 1e-9, 1e-9, 1e-9, 16, 0)
    assert lib.mmx_create(ctypes.byref(h), ctypes.byref(d)) == L.EINVAL
