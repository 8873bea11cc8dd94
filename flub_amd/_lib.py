"""ctypes binding of libflub.so (the C-ABI declared in include/flub.h).

The library is the only compute path: if it is missing this module raises
instead of falling back to anything else.  Build it with ``python -m
flub_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libflub.so")

_P = ctypes.c_void_p
_I = ctypes.c_int
_Z = ctypes.c_size_t
_D = ctypes.c_double
_F = ctypes.c_float
_PP = ctypes.POINTER(ctypes.c_void_p)
_PI = ctypes.POINTER(ctypes.c_int)
_PD = ctypes.POINTER(ctypes.c_double)
_PF = ctypes.POINTER(ctypes.c_float)

# name -> argtypes (every function returns int unless listed in _RESTYPE)
SIGNATURES = {
    "flub_last_error": [],
    "flub_device_count": [_PI],
    "flub_malloc": [_PP, _Z],
    "flub_free": [_P],
    "flub_memcpy_h2d": [_P, _P, _Z, _P],
    "flub_memcpy_d2h": [_P, _P, _Z, _P],
    "flub_memcpy_d2d": [_P, _P, _Z, _P],
    "flub_memset": [_P, _I, _Z, _P],
    "flub_fill_f32": [_P, _F, _Z, _P],
    "flub_fill_i32": [_P, ctypes.c_int32, _Z, _P],
    "flub_stream_create": [_PP],
    "flub_stream_destroy": [_P],
    "flub_stream_synchronize": [_P],
    "flub_device_synchronize": [],
    "flub_event_create": [_PP],
    "flub_event_destroy": [_P],
    "flub_event_record": [_P, _P],
    "flub_event_elapsed_ms": [_P, _P, _PF],
    "flub_tag_dirichlet": [_P, _P, _Z, _P],
    "flub_tag_neumann": [_P, _P, _Z, _P],
    "flub_divergence_tag": [_P] * 8 + [_I, _I, _I, _P],
    "flub_phi_prims": [_P, _P, _I, _I, _I, _F, _F, _F, _F, _I, _P, _P, _P, _P, _P],
    "flub_phi_to_fraction": [_P, _P, _I, _I, _I, _I, _P],
    "flub_face_fractions": [_P] * 6 + [_I, _I, _I, _F, _I, _P, _P, _P, _P, _P],
    "flub_update_argmin": [_P, _P, _I, _I, _I, _P, _I, _I, _I, _I, _P, _P, _F, _P, _P, _F, _P],
    "flub_solid_vel": [_P, _P, _I, _I, _I, _I, _F, _P, _P, _P, _P],
    "flub_divergence_fractional_3d": [_P] * 10 + [_I, _I, _I, _P],
    "flub_set_unknown": [_P, _P, _P, _P, _I, _I, _I, _P],
    "flub_advect_mc": [_P, _P, _P, _I, _I, _I, _P, _P, _I, _I, _I, _F, _F, _F, _F, _F, _P],
    "flub_advect_phi": [_P, _P, _P, _I, _I, _I, _P, _P, _F, _F, _F, _F, _F, _P],
    "flub_advect_simple_3d": [_P] * 5 + [_I, _I, _I, _F, _F, _P],
    "flub_extrapolate_3d": [_P, _P, _I, _I, _I, _P],
    "flub_flush_1_3d": [_P, _Z, _P],
    "flub_extrapolate_by_1": [_P] * 6 + [_I, _I, _I, _P],
    "flub_extrapolate": [_P] * 7 + [_I, _I, _I, _I, _P],
    "flub_mask_with_tag": [_P, _P, _Z, _P],
    "flub_scale": [_P, _F, _Z, _P],
    "flub_subtract_gradient_tag": [_P, _P, _P, _P, _I, _I, _I, _I, _P],
    "flub_subtract_gradient_tag_f64": [_P, _P, _P, _P, _I, _I, _I, _I, _P],
    "flub_inner": [_P, _P, _Z, _P, _P],
    "flub_inner_noalloc": [_P, _P, _Z, _P, _P],
    "flub_mgpcg_create": [_PP, _I, _I, _I, _I, _I, _P],
    "flub_mgpcg_destroy": [_P],
    "flub_mgpcg_set_smoothing": [_P, _I, _I],
    "flub_mgpcg_set_coarse_scale": [_P, _F],
    "flub_mgpcg_set_sweep_cap": [_P, _I],
    "flub_mgpcg_set_tags": [_P, _P],
    "flub_mgpcg_init": [_P, _P, _F],
    "flub_mgpcg_mask_interior": [_P],
    "flub_mgpcg_zero_centered": [_P],
    "flub_mgpcg_apply_preconditioner": [_P, _I],
    "flub_mgpcg_graph_precondition": [_P, _I],
    "flub_mgpcg_compute_Ap": [_P, _I],
    "flub_mgpcg_update_p": [_P],
    "flub_mgpcg_update_r": [_P],
    "flub_mgpcg_update_x": [_P],
    "flub_mgpcg_update_alpha": [_P, _D],
    "flub_mgpcg_update_beta": [_P, _D],
    "flub_mgpcg_inner": [_P, _I, _I, _I],
    "flub_mgpcg_solve": [_P, _I, _D, _D, _D, _I, _I, _PI, _PD],
    "flub_mgpcg_buffer": [_P, _I, _I, _PP],
    "flub_mgpcg_compute_divd": [_P, _I, _P],
    "flub_mgpcg_scalars": [_P, _PD],
    "flub_mgpcg_level_shape": [_P, _I, _PI],
    "flub_mgpcg_profile": [_P, _I, _I, _PF, _PD],
    "flub_mgpcg_probe": [_P, _I, _I],
    "flub_mgpcg_probe_read": [_P, _P, _P, _P],
    "flub_mgpcg_apply_A": [_P, _P, _P],
    "flub_mgpcg_residual_norm": [_P, _P, _P, _PD],
    "flub_mgpcg_true_residual_f64": [_P, _P, _P, _P, _PD],
    "flub_mgpcg_solution_f64": [_P, _P],
    "flub_mgpcg_precondition_into": [_P, _P, _P, _I],
    "flub_project": [_P] * 8 + [_F, _I, _D, _D, _I, _P, _P, _PI, _PD],
    "flub_conv3d_bf16": [_P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _I, _P],
    "flub_conv3d_wgrad_bf16": [_P, _P, _P, _P, _Z, _I, _I, _I, _I, _I, _I, _P],
    "flub_head_fwd": [_P, _P, _F, _P, _Z, _I, _P],
    "flub_head_fwd_masked": [_P, _P, _F, _P, _Z, _I, _P, _F, _P],
    "flub_conv3d_upcat_bf16": [_P, _I, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _I, _P],
    "flub_pool2_bf16": [_P, _P, _I, _I, _I, _I, _I, _P],
    "flub_pack_input_bf16": [_P, _P, _F, _P, _Z, _P],
}
_RESTYPE = {"flub_last_error": ctypes.c_char_p}

_lib = None


class FlubError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, name, code, msg):
        super().__init__(f"{name} failed with code {code}: {msg}")
        self.code = code


def lib():
    """Load libflub.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        # One HIP runtime per process: PyTorch-ROCm ships its own libamdhip64
        # (same soname as the system one libflub links).  Loading PyTorch first
        # makes libflub bind to that already-loaded runtime, so both share one
        # device context; loading libflub first would bring in a second runtime
        # and PyTorch would then find no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build the HIP library first (python -m flub_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPE.get(name, ctypes.c_int)
        _lib = L
    return _lib


def last_error() -> str:
    s = lib().flub_last_error()
    return s.decode() if s else ""


def call(name, *args, allow=()):
    """Call a C-ABI function; raise FlubError on a status outside 0/allow."""
    rc = getattr(lib(), name)(*args)
    if rc != 0 and rc not in allow:
        raise FlubError(name, rc, last_error())
    return rc


# Status codes (csrc/common.h)
OK = 0
EINVAL = -1
ESHAPE = -2
ESTATE = -3
ENOTCONV = -4
ENONFINITE = -5
