"""The C-ABI library: it loads without a GPU, exports every function that
include/flub.h declares, the ctypes table matches the header, and argument
errors are reported (as status codes with text) before any device work."""
import ctypes
import os
import re

import pytest

from flub_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    hdr = open(os.path.join(ROOT, "include", "flub.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return set(re.findall(r"\b(flub_[A-Za-z0-9_]+)\s*\(", hdr))


def test_library_exports_every_declared_function():
    L = _lib.lib()
    names = header_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    assert set(_lib.SIGNATURES) == header_functions()


def test_argument_errors_without_device_work():
    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.flub_mgpcg_create(ctypes.byref(h), 64, 64, 64, 0, 0, None) == _lib.EINVAL
    assert "n_mg_levels" in _lib.last_error()
    assert L.flub_mgpcg_create(ctypes.byref(h), 64, 64, 64, 4, 7, None) == _lib.EINVAL
    assert L.flub_mgpcg_create(ctypes.byref(h), 64, 64, 6, 3, 0, None) == _lib.ESHAPE     # 6 >> 2 = 1: odd z
    assert "even" in _lib.last_error()
    assert L.flub_mgpcg_create(ctypes.byref(h), 2000, 2000, 2000, 3, 0, None) == _lib.ESHAPE
    assert L.flub_mgpcg_create(None, 8, 8, 8, 2, 0, None) == _lib.EINVAL
    assert L.flub_mgpcg_solve(None, 1, 0.0, 0.0, 0.0, 1, 0, None, None) == _lib.EINVAL
    assert L.flub_subtract_gradient_tag(1, 1, 1, 1, 4, 4, 4, 4, None) == _lib.EINVAL      # d must be 1..3
    assert L.flub_divergence_tag(None, 1, 1, 1, 1, 1, 1, 1, 4, 4, 4, None) == _lib.EINVAL
    assert L.flub_divergence_tag(1, 1, 1, 1, 1, 1, 1, 1, 0, 4, 4, None) == _lib.ESHAPE
    assert L.flub_malloc(None, 16) == _lib.EINVAL
    p = ctypes.c_void_p(123)
    assert L.flub_malloc(ctypes.byref(p), 0) == 0 and p.value is None


def test_python_errors_are_raised():
    with pytest.raises(_lib.FlubError) as e:
        _lib.call("flub_mgpcg_create", None, 8, 8, 8, 2, 0, None)
    assert e.value.code == _lib.EINVAL
