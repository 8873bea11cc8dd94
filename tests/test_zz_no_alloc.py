"""flub_inner_noalloc: the inner product without any allocation (its FP64
partials live in a static device array) against flub_inner through the C-ABI:
bitwise the same value on two calls in a row, the second over fewer elements
than the first (the array keeps the first call's partials), and sizes beyond
the stated maximum rejected before any device work."""
import ctypes

import numpy as np
import pytest
import torch

from flub_amd import _lib

pytestmark = pytest.mark.gpu


def _inner(name, a, b, n):
    out = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
    _lib.call(name, ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()), ctypes.c_size_t(n),
              ctypes.c_void_p(out.data_ptr()), None)
    torch.cuda.synchronize()
    return float(out.cpu()[0])


def test_noalloc_inner_equals_inner_on_consecutive_calls():
    rng = np.random.default_rng(41)
    for n in (5_000_003, 100_003, 0):
        a = rng.normal(size=max(n, 1)).astype(np.float32)
        b = rng.normal(size=max(n, 1)).astype(np.float32)
        da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
        x = _inner("flub_inner_noalloc", da, db, n)
        y = _inner("flub_inner", da, db, n)
        assert np.float64(x).view(np.uint64) == np.float64(y).view(np.uint64), (n, x, y)
        ref = float(np.dot(a[:n].astype(np.float64), b[:n].astype(np.float64)))
        assert abs(x - ref) <= 1e-12 * max(1.0, float(np.sum(np.abs(a[:n].astype(np.float64) * b[:n])))), (n, x, ref)


def test_noalloc_inner_rejects_sizes_beyond_the_maximum():
    d = torch.zeros(4, dtype=torch.float32, device="cuda")
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    L = _lib.lib()
    rc = L.flub_inner_noalloc(ctypes.c_void_p(d.data_ptr()), ctypes.c_void_p(d.data_ptr()),
                              ctypes.c_size_t(2 ** 31), ctypes.c_void_p(out.data_ptr()), None)
    assert rc == _lib.ESHAPE
    assert "exceeds the maximum" in _lib.last_error()
