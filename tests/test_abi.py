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


def test_probe_library_rejects_shapes_before_any_device_call():
    """libfomo_probe.so (bench.py --pattern-probe) checks its arguments before
    it allocates or launches anything, so these calls never reach HIP."""
    import ctypes
    from fomo_julia_amd import build as B
    if not os.path.exists(B.PROBE_PATH):
        pytest.skip("libfomo_probe.so (measurement only) is not built")
    lib = ctypes.CDLL(B.PROBE_PATH)
    lib.fomo_probe_pattern.argtypes = [ctypes.c_int] * 9 + [ctypes.POINTER(ctypes.c_double)]
    lib.fomo_probe_stream.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_double)]
    us = ctypes.c_double()
    ok = dict(nread=8, nwrite=5, cpt=2, nx=4608, nz=3920, tz=40, grid=1024, lds=0, reps=20)
    bad = [dict(nread=7), dict(nwrite=4), dict(cpt=3), dict(nx=4600), dict(nz=3921), dict(tz=1),
           dict(grid=0), dict(lds=65537), dict(lds=-1), dict(reps=0)]
    for change in bad:
        a = {**ok, **change}
        assert lib.fomo_probe_pattern(*a.values(), ctypes.byref(us)) == -1, change
    assert lib.fomo_probe_pattern(*ok.values(), None) == -1
    for mode, nbytes, grid in ((3, 1 << 20, 8), (0, 0, 8), (0, 100, 8), (1, 1 << 20, 0)):
        assert lib.fomo_probe_stream(mode, nbytes, grid, 20, ctypes.byref(us)) == -1
    lib.fomo_probe_runtime.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_double)]
    for mode, nbytes, reps in ((3, 1 << 20, 20), (-1, 1 << 20, 20), (0, 0, 20), (1, 100, 20),
                               (2, 1 << 20, 0)):
        assert lib.fomo_probe_runtime(mode, nbytes, reps, ctypes.byref(us)) == -1
    assert lib.fomo_probe_runtime(0, 1 << 20, 20, None) == -1
