"""Device arrays on hipMalloc memory, with the subset of the Warp array API the
reference's projection path uses (const_param.py:114-150, multigrid/mgpcg.py:
255-268): ``empty``, ``zeros``, ``zeros_like``, ``from_numpy``, ``copy``,
``.numpy()``, ``.zero_()``, ``.fill_()``, ``.shape``.

Python ``float`` means float32 and ``int`` means int32, as in Warp.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

_DTYPES = {float: np.float32, int: np.int32}


def _np_dtype(dtype):
    if dtype in _DTYPES:
        return np.dtype(_DTYPES[dtype])
    return np.dtype(dtype)


class DeviceArray:
    """A C-contiguous array in device memory (k fastest for 3-D fields)."""

    def __init__(self, shape, dtype=float, ptr=None, owner=None):
        if isinstance(shape, int):
            shape = (shape,)
        self.shape = tuple(int(s) for s in shape)
        self.dtype = _np_dtype(dtype)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * self.dtype.itemsize
        self._owned = ptr is None
        self._owner = owner  # keeps a parent alive for views
        if ptr is None:
            p = ctypes.c_void_p()
            _lib.call("flub_malloc", ctypes.byref(p), self.nbytes)
            ptr = p.value
        self.ptr = int(ptr or 0)

    # -- Warp-like API
    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _lib.call("flub_memcpy_d2h", out.ctypes.data, self.ptr, self.nbytes, None)
        return out

    def assign(self, arr) -> "DeviceArray":
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        if a.size != self.size:
            raise ValueError(f"size mismatch: {a.shape} vs {self.shape}")
        if self.nbytes:
            _lib.call("flub_memcpy_h2d", self.ptr, a.ctypes.data, self.nbytes, None)
        return self

    def zero_(self) -> "DeviceArray":
        _lib.call("flub_memset", self.ptr, 0, self.nbytes, None)
        return self

    def fill_(self, value) -> "DeviceArray":
        if self.dtype == np.float32:
            _lib.call("flub_fill_f32", self.ptr, float(value), self.size, None)
        elif self.dtype == np.int32:
            _lib.call("flub_fill_i32", self.ptr, int(value), self.size, None)
        else:
            self.assign(np.full(self.shape, value, self.dtype))
        return self

    def __len__(self):
        return self.shape[0]

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype.name}, ptr=0x{self.ptr:x})"

    def __del__(self):
        if getattr(self, "_owned", False) and self.ptr and _lib is not None:
            try:
                _lib.lib().flub_free(self.ptr)
            except Exception:
                pass
            self.ptr = 0


def empty(shape, dtype=float) -> DeviceArray:
    return DeviceArray(shape, dtype)


def zeros(shape, dtype=float) -> DeviceArray:
    return DeviceArray(shape, dtype).zero_()


def zeros_like(a: DeviceArray) -> DeviceArray:
    return zeros(a.shape, a.dtype)


def from_numpy(arr, dtype=None, shape=None) -> DeviceArray:
    a = np.asarray(arr)
    dt = _np_dtype(dtype) if dtype is not None else a.dtype
    if dt == np.float64:
        dt = np.dtype(np.float32)
    if dt == np.int64:
        dt = np.dtype(np.int32)
    a = np.ascontiguousarray(a, dtype=dt)
    if shape is not None:
        a = a.reshape(shape)
    return DeviceArray(a.shape, dt).assign(a)


def copy(dest: DeviceArray, src: DeviceArray) -> None:
    """wp.copy: dest <- src (same byte size)."""
    if dest.nbytes != src.nbytes:
        raise ValueError(f"copy size mismatch: {dest.shape}/{dest.dtype} vs {src.shape}/{src.dtype}")
    _lib.call("flub_memcpy_d2d", dest.ptr, src.ptr, src.nbytes, None)


def synchronize() -> None:
    _lib.call("flub_device_synchronize")


def device_count() -> int:
    c = ctypes.c_int(0)
    _lib.call("flub_device_count", ctypes.byref(c))
    return c.value


class ScopedTimer:
    """wp.ScopedTimer(name, synchronize=True) on HIP events: elapsed device time
    of the enclosed stream-0 work, in ms (``.elapsed``); optionally stored in
    ``dict[name]`` as the reference's timing dicts are (example_fluid.py:130)."""

    def __init__(self, name, active=True, synchronize=True, dict=None, verbose=False):
        self.name, self.active, self.dict, self.verbose = name, active, dict, verbose
        self.elapsed = 0.0

    def __enter__(self):
        if self.active:
            self._a, self._b = ctypes.c_void_p(), ctypes.c_void_p()
            _lib.call("flub_event_create", ctypes.byref(self._a))
            _lib.call("flub_event_create", ctypes.byref(self._b))
            _lib.call("flub_device_synchronize")
            _lib.call("flub_event_record", self._a, None)
        return self

    def __exit__(self, *exc):
        if self.active:
            _lib.call("flub_event_record", self._b, None)
            _lib.call("flub_device_synchronize")
            ms = ctypes.c_float()
            _lib.call("flub_event_elapsed_ms", self._a, self._b, ctypes.byref(ms))
            self.elapsed = float(ms.value)
            _lib.call("flub_event_destroy", self._a)
            _lib.call("flub_event_destroy", self._b)
            if self.dict is not None:
                self.dict[self.name] = self.elapsed
            if self.verbose:
                print(f"{self.name} took {self.elapsed:.3f} ms")
        return False
