"""Rigid bodies as the projection sees them (rigid_coupling.py of the
reference): their poses and velocities, and ``precompute_solid_vel``
(rigid_coupling.py:51-70), which sets the rigid face velocities that
divergence_tag / divergence_fractional_3d and subtract_gradient_tag read.

``RigidBodies`` carries the attributes the projection path reads from the
reference's ``RigidCoupling`` (``x_pair``, ``theta_pair``, ``v``, ``omega``,
``rigid_geometry``, ``cached_levelset``, ``precompute_solid_vel_3d``); the
rigid-body dynamics and scene loading of that class are outside this path.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from . import const_param as cp
from . import device as wp_dev
from .geometry import GeometryArray, GeometryType3D
from .launch import Kernel, launch
from .utils.util import TexPair


def _precompute_solid_vel(x, v, omega, vel_rigid, vel_owner, d, h0=None):
    uf = cp.get_uf(vel_rigid, d).f
    owner = cp.get_uf(vel_owner, d).f
    if owner.shape != uf.shape:
        raise ValueError("precompute_solid_vel: owner and velocity arrays differ in shape")
    if not (x.size == v.size == omega.size):
        raise ValueError("precompute_solid_vel: body arrays differ in length")
    h0 = cp.h0 if h0 is None else h0
    _lib.call("flub_solid_vel", uf.ptr, owner.ptr, *uf.shape, int(d), float(h0), x.ptr, v.ptr, omega.ptr, None)


precompute_solid_vel = Kernel("precompute_solid_vel", _precompute_solid_vel,
                              lambda x, v, om, vel, own, d, h0=None: cp.get_uf(vel, d).f.shape)


class RigidBodies:
    """n rigid bodies: positions x (n, 3), orientations q (n, 4) as (x, y, z, w),
    linear velocities v and angular velocities omega (n, 3), and their
    GeometryType3D list; all held in device memory."""

    def __init__(self, x, q, geometry, v=None, omega=None, h0=None):
        geometry = [g if isinstance(g, GeometryType3D) else GeometryType3D(*g) for g in geometry]
        n = len(geometry)
        x = np.asarray(x, np.float32).reshape(n, 3)
        q = np.asarray(q, np.float32).reshape(n, 4)
        self.x_pair = TexPair(wp_dev.from_numpy(x), wp_dev.from_numpy(x))
        self.theta_pair = TexPair(wp_dev.from_numpy(q), wp_dev.from_numpy(q))
        zero = np.zeros((n, 3), np.float32)
        self.v = wp_dev.from_numpy(zero if v is None else np.asarray(v, np.float32).reshape(n, 3))
        self.omega = wp_dev.from_numpy(zero if omega is None else np.asarray(omega, np.float32).reshape(n, 3))
        self.rigid_geometry = GeometryArray(geometry)
        self.cached_levelset = {}
        self.h0 = h0   # grid spacing; None: the module grid's h0 (const_param.h0)

    def precompute_solid_vel_3d(self, x, theta, v, omega, geometry, vel_rigid, vel_owner):
        """rigid_coupling.py:225-228."""
        for d, a in ((1, vel_rigid.u), (2, vel_rigid.v), (3, vel_rigid.w)):
            launch(precompute_solid_vel, dim=a.f.shape, inputs=[x, v, omega, vel_rigid, vel_owner, d, self.h0])
