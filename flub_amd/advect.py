"""Advection (advect.py:11-137 of the reference) on the HIP C-ABI:

    advect_mc(vel, q_in, q_out)          # MacCormack, one field (advect.py:133-137)
    advect_uv_mc(vel_in, vel_out)        # the three velocity components
    advect_3d(vel, qf, new_vel, new_qf)  # a scalar field and the velocity
    launch(advect_phi, dim=shape, inputs=[vel, phi, new_phi])
    launch(advect_simple_3d, dim=shape, inputs=[u, v, w, q, new_q, dt, h])

As in the reference, advect_mc and advect_phi take the time step and cell size
from the module grid (const_param.dt, const_param.h0); both can be given
explicitly.
"""
from __future__ import annotations

from . import _lib
from . import const_param as cp
from .launch import Kernel


def _grid_of(vel):
    nx, ny, nz = vel.v.f.shape[0], vel.u.f.shape[1], vel.u.f.shape[2]
    for d in (1, 2, 3):
        if cp.get_uf(vel, d).f.shape != cp.field_shape(d, (nx, ny, nz)):
            raise ValueError("velocity arrays do not form a MAC grid")
    return nx, ny, nz


def advect_mc(args_common, q_in, q_out, dt=None, h0=None):
    """advect.py:133-137 (the three MacCormack launches, fused)."""
    nx, ny, nz = _grid_of(args_common)
    if q_in.f.shape != q_out.f.shape:
        raise ValueError("advect_mc: q_in and q_out differ in shape")
    v = args_common
    _lib.call("flub_advect_mc", v.u.f.ptr, v.v.f.ptr, v.w.f.ptr, nx, ny, nz, q_in.f.ptr, q_out.f.ptr, *q_in.f.shape,
              *(float(x) for x in q_in.stagger), float(cp.dt if dt is None else dt),
              float(cp.h0 if h0 is None else h0), None)


def advect_uv_mc(vel_in, vel_out, dt=None, h0=None):
    """advect.py:139-142."""
    advect_mc(vel_in, vel_in.u, vel_out.u, dt, h0)
    advect_mc(vel_in, vel_in.v, vel_out.v, dt, h0)
    advect_mc(vel_in, vel_in.w, vel_out.w, dt, h0)


def advect_3d(vel, qf, new_vel, new_qf, dt=None, h0=None):
    """advect.py:145-147."""
    advect_mc(vel, qf, new_qf, dt, h0)
    advect_uv_mc(vel, new_vel, dt, h0)


def _advect_phi(vel, phi, new_phi, dt=None, h0=None):
    nx, ny, nz = _grid_of(vel)
    if phi.f.shape != (nx, ny, nz) or new_phi.f.shape != (nx, ny, nz):
        raise ValueError("advect_phi: level set does not match the grid")
    _lib.call("flub_advect_phi", vel.u.f.ptr, vel.v.f.ptr, vel.w.f.ptr, nx, ny, nz, phi.f.ptr, new_phi.f.ptr,
              *(float(x) for x in phi.stagger), float(cp.dt if dt is None else dt),
              float(cp.h0 if h0 is None else h0), None)


def _advect_simple_3d(uf, vf, wf, qf, new_qf, dt, h):
    nx, ny, nz = qf.shape
    if uf.shape != (nx + 1, ny, nz) or vf.shape != (nx, ny + 1, nz) or wf.shape != (nx, ny, nz + 1):
        raise ValueError("advect_simple_3d: velocity arrays do not match q")
    _lib.call("flub_advect_simple_3d", uf.ptr, vf.ptr, wf.ptr, qf.ptr, new_qf.ptr, nx, ny, nz, float(dt), float(h),
              None)


advect_phi = Kernel("advect_phi", _advect_phi, lambda vel, phi, new_phi, *a: phi.f.shape)
advect_simple_3d = Kernel("advect_simple_3d", _advect_simple_3d, lambda uf, vf, wf, qf, *a: qf.shape)
