"""flub_amd -- MI355X pressure projection (divergence -> MGPCG -> gradient
subtraction) behind the reference's Python call signatures, on hand-written
HIP kernels reached through the C-ABI in include/flub.h (libflub.so)."""
from . import _lib  # noqa: F401
from .device import DeviceArray, copy, empty, from_numpy, zeros, zeros_like  # noqa: F401

__all__ = ["DeviceArray", "copy", "empty", "from_numpy", "zeros", "zeros_like"]
