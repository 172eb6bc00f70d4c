"""The C-ABI library loads and exports every symbol include/fomo_mi355x.h
declares (no compute: this runs without a GPU)."""
import os
import re

import pytest

from fomo_julia_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "fomo_mi355x.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fomo_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    names = declared_functions()
    assert len(names) >= 30
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libfomo_mi355x.so not built (python -m fomo_julia_amd.build)")
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_errors_without_device_are_reported_not_crashing():
    lib = _lib.load()
    import ctypes
    h = ctypes.c_void_p()
    rc = lib.fomo_ctx_create(-1, ctypes.byref(h))
    assert rc < 0
    assert lib.fomo_ctx_last_error(None)
    assert lib.fomo_ctx_set_tiling(None, 2, 32, 16) == -1
