"""Velocity extrapolation into the air (extrapolate.py:9-47 of the reference)
and its driver (water_3d.py:70-73), on the HIP C-ABI:

    launch(set_unknown, dim=shape, inputs=[phi, vel_owner])      # utils/util.py:424-437
    extrapolate_by_1(vel, vel_owner)                             # extrapolate.py:40-47
    launch(extrapolate_3d, dim=shapeu, inputs=[vel, tags, 1])    # one axis, one launch
    launch(flush_1_3d, dim=shapeu, inputs=[tags.u.f])

The reference's flush turns every -1 mark into 1: the marks of extrapolated
faces, and the -1 owner values that the last face layer of the reused owner
arrays still holds (set_unknown writes only the first nx x ny x nz entries).
After two passes every pass recomputes the same faces from the same known
values, so ``extrapolate(..., n_cells)`` runs min(n_cells, 2) passes, which
gives the fields of n_cells passes.
"""
from __future__ import annotations

from . import _lib
from . import const_param as cp
from .launch import Kernel


def _check_mac(vel, tags):
    for d in (1, 2, 3):
        if cp.get_uf(vel, d).f.shape != cp.get_uf(tags, d).f.shape:
            raise ValueError("velocity and mark arrays differ in shape")
    nx, ny, nz = vel.v.f.shape[0], vel.u.f.shape[1], vel.u.f.shape[2]
    for d in (1, 2, 3):
        if cp.get_uf(vel, d).f.shape != cp.field_shape(d, (nx, ny, nz)):
            raise ValueError("face arrays do not form a MAC grid")
    return nx, ny, nz


def _extrapolate_3d(mac, u, d):
    qf = cp.get_uf(mac, d).f
    unknown = cp.get_uf(u, d).f
    if qf.shape != unknown.shape:
        raise ValueError("extrapolate_3d: velocity and mark arrays differ in shape")
    _lib.call("flub_extrapolate_3d", qf.ptr, unknown.ptr, *qf.shape, None)


def _flush_1_3d(tag):
    _lib.call("flub_flush_1_3d", tag.ptr, tag.size, None)


extrapolate_3d = Kernel("extrapolate_3d", _extrapolate_3d, lambda mac, u, d: cp.get_uf(mac, d).f.shape)
flush_1_3d = Kernel("flush_1_3d", _flush_1_3d, lambda tag: tag.shape)


def extrapolate_by_1(vel, tag):
    """extrapolate.py:40-47 (each axis' extrapolate + flush as one kernel)."""
    nx, ny, nz = _check_mac(vel, tag)
    _lib.call("flub_extrapolate_by_1", vel.u.f.ptr, vel.v.f.ptr, vel.w.f.ptr, tag.u.f.ptr, tag.v.f.ptr, tag.w.f.ptr,
              nx, ny, nz, None)


def extrapolate(vel, vel_owner, phi, n_cells=None):
    """water_3d.py:70-73: set_unknown from the water level set phi (a Field),
    then n_cells (default 16 * res / 128) passes of extrapolate_by_1."""
    nx, ny, nz = _check_mac(vel, vel_owner)
    if phi.f.shape != (nx, ny, nz):
        raise ValueError("extrapolate: phi does not match the grid")
    n = 16 * nx // 128 if n_cells is None else int(n_cells)
    _lib.call("flub_extrapolate", vel.u.f.ptr, vel.v.f.ptr, vel.w.f.ptr, vel_owner.u.f.ptr, vel_owner.v.f.ptr,
              vel_owner.w.f.ptr, phi.f.ptr, nx, ny, nz, n, None)
