"""MAC-grid divergence (divergence.py:5-41 ``divergence_fractional_3d``,
:44-99 ``divergence_tag`` and :174-180 ``mask_with_tag`` of the reference),
launched on the HIP C-ABI:

    launch(divergence_tag, dim=shape, inputs=[vel, div, vel_rigid, tag])
    launch(divergence_fractional_3d, dim=shape,
           inputs=[u, v, w, div, fu, fv, fw, u_rigid, v_rigid, w_rigid])
    launch(mask_with_tag, dim=shape, inputs=[div, tag])

(divergence_fractional_3d takes the bare arrays, as the reference's kernel does.)
"""
from __future__ import annotations

from . import _lib
from .launch import Kernel


def _divergence_tag(vels, velocity_divs, vel_rigid, tag):
    nx, ny, nz = tag.f.shape
    _check(vels, vel_rigid, velocity_divs, tag)
    _lib.call("flub_divergence_tag", vels.u.f.ptr, vels.v.f.ptr, vels.w.f.ptr, velocity_divs.f.ptr,
              vel_rigid.u.f.ptr, vel_rigid.v.f.ptr, vel_rigid.w.f.ptr, tag.f.ptr, nx, ny, nz, None)


def _check(vels, vel_rigid, div, tag):
    nx, ny, nz = tag.f.shape
    want = {"u": (nx + 1, ny, nz), "v": (nx, ny + 1, nz), "w": (nx, ny, nz + 1)}
    for g in (vels, vel_rigid):
        for c, s in want.items():
            if getattr(g, c).f.shape != s:
                raise ValueError(f"face field {c} has shape {getattr(g, c).f.shape}, expected {s}")
    if div.f.shape != (nx, ny, nz):
        raise ValueError("divergence field shape must equal the tag shape")
    if tag.f.dtype.itemsize != 4 or div.f.dtype.itemsize != 4:
        raise TypeError("tag must be int32 and div float32")


def _mask_with_tag(div, tag):
    if div.f.shape != tag.f.shape:
        raise ValueError("mask_with_tag: shape mismatch")
    _lib.call("flub_mask_with_tag", div.f.ptr, tag.f.ptr, div.f.size, None)


divergence_tag = Kernel("divergence_tag", _divergence_tag, lambda vels, div, vr, tag: tag.f.shape)
mask_with_tag = Kernel("mask_with_tag", _mask_with_tag, lambda div, tag: div.f.shape)


def _divergence_fractional_3d(uf, vf, wf, velocity_divs, fu, fv, fw, u_rigid, v_rigid, w_rigid):
    nx, ny, nz = velocity_divs.shape
    want = {"u": (nx + 1, ny, nz), "v": (nx, ny + 1, nz), "w": (nx, ny, nz + 1)}
    for name, arrs in (("u", (uf, fu, u_rigid)), ("v", (vf, fv, v_rigid)), ("w", (wf, fw, w_rigid))):
        for a in arrs:
            if a.shape != want[name]:
                raise ValueError(f"divergence_fractional_3d: {name} array has shape {a.shape}, expected {want[name]}")
    _lib.call("flub_divergence_fractional_3d", uf.ptr, vf.ptr, wf.ptr, velocity_divs.ptr, fu.ptr, fv.ptr, fw.ptr,
              u_rigid.ptr, v_rigid.ptr, w_rigid.ptr, nx, ny, nz, None)


divergence_fractional_3d = Kernel("divergence_fractional_3d", _divergence_fractional_3d,
                                  lambda u, v, w, div, *rest: div.shape)
