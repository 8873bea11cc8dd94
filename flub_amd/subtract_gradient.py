"""Pressure-gradient subtraction (subtract_gradient.py:38-68
``subtract_gradient_tag`` of the reference) on the HIP C-ABI:

    launch(subtract_gradient_tag, dim=shapeu, inputs=[vel, vel_rigid, pf, tag, 1])  # and 2, 3
"""
from __future__ import annotations

from . import _lib
from .const_param import get_uf
from .launch import Kernel


def _subtract_gradient_tag(vel, vel_rigid, pf, tag, d):
    nx, ny, nz = tag.f.shape
    uf = get_uf(vel, d).f
    ur = get_uf(vel_rigid, d).f
    want = (nx + (d == 1), ny + (d == 2), nz + (d == 3))
    if uf.shape != want or ur.shape != want or pf.f.shape != (nx, ny, nz):
        raise ValueError(f"subtract_gradient_tag: face shape {uf.shape} / pressure {pf.f.shape} "
                         f"do not match tag {tag.f.shape} for d={d}")
    _lib.call("flub_subtract_gradient_tag", uf.ptr, ur.ptr, pf.f.ptr, tag.f.ptr, nx, ny, nz, int(d), None)


def _dim(vel, vel_rigid, pf, tag, d):
    return get_uf(vel, d).f.shape


subtract_gradient_tag = Kernel("subtract_gradient_tag", _subtract_gradient_tag, _dim)
