"""Cell classification and solid geometry of the reference (geometry.py), on
the HIP C-ABI:

* ``_tag_dirichlet`` (geometry.py:335-344) and ``_tag_neumann`` (:390-397), in
  the order example_fluid.py:111-114 applies them: water phi -> AIR/FLUID,
  then rigid phi < 0 -> SOLID;
* ``GeometryType3D`` (geometry.py:8-13) and the primitive signed distances
  (:17-32 ``_sdf_cube``, :82-113 ``compute_sdf_atom_prims``, :148-169
  ``sdf_argmin_prims``);
* ``precompute_grid_point_phi_prims`` (:225-239) and ``phi_to_fraction``
  (:264-290) as launchable kernels;
* ``update_argmin`` (:365-388): a body given as a cached level set (a mesh
  body's signed distance on a grid of its own), resampled tricubically in the
  body frame into a distance array;
* ``PrecomputeFaceFraction`` (:401-462): tags, rigid level sets and face
  fractions with their owners, primitive bodies and cached level sets
  (``rc.cached_levelset``: body index -> (level set, ijk, h_levelset)).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib
from . import const_param as cp
from . import device as wp_dev
from .launch import Kernel, launch

# body kinds of GeometryType3D.geometry (geometry.py:34)
GEOM_NONE, GEOM_QUAD, GEOM_CIRCLE, GEOM_RECT, GEOM_X, GEOM_MESH = range(6)


@dataclass
class GeometryType3D:
    """geometry.py:8-13: kind, half-extents (radius[0] for a circle), density,
    mesh id and volume of one rigid body."""
    geometry: int = GEOM_NONE
    radius: tuple = (0.0, 0.0, 0.0)
    rho: float = 1.0
    mesh_id: int = 0
    volume: float = 0.0


class GeometryArray:
    """The device form of a list of GeometryType3D: kind (int32) and radius
    (float32 x 3) arrays, as the kernels read them."""

    def __init__(self, geoms):
        geoms = list(geoms)
        self.items = geoms
        self.gtype = wp_dev.from_numpy(np.array([g.geometry for g in geoms], np.int32).reshape(-1))
        self.radius = wp_dev.from_numpy(np.array([tuple(g.radius) for g in geoms], np.float32).reshape(-1, 3))

    @property
    def shape(self):
        return (len(self.items),)


def _tag_dirichlet(tag, phi):
    if tag.f.shape != phi.f.shape:
        raise ValueError("tag_dirichlet: shape mismatch")
    _lib.call("flub_tag_dirichlet", phi.f.ptr, tag.f.ptr, tag.f.size, None)


def _tag_neumann(tag, phi):
    if tag.f.shape != phi.f.shape:
        raise ValueError("tag_neumann: shape mismatch")
    _lib.call("flub_tag_neumann", phi.f.ptr, tag.f.ptr, tag.f.size, None)


tag_dirichlet_kernel = Kernel("_tag_dirichlet", _tag_dirichlet, lambda tag, phi: tag.f.shape)
tag_neumann_kernel = Kernel("_tag_neumann", _tag_neumann, lambda tag, phi: tag.f.shape)


def tag_dirichlet(tag, phi):
    _tag_dirichlet(tag, phi)


def tag_neumann(tag, phi_rigid):
    _tag_neumann(tag, phi_rigid)


def _bodies(x, q, geometry):
    n = geometry.shape[0]
    if x.size != 3 * n or q.size != 4 * n:
        raise ValueError("rigid arrays: x must hold 3 and q 4 floats per body")
    return n, x.ptr, q.ptr, geometry.gtype.ptr, geometry.radius.ptr


def _phi_prims(phi, owner, x, q, geometry, stagger):
    """geometry.py:225-239 (owner written only for a staggered array, as there)."""
    n, xp, qp, gp, rp = _bodies(x, q, geometry)
    staggered = any(float(s) != 0.0 for s in stagger)
    own = owner.ptr if (staggered and owner is not None and owner.size == phi.size) else None
    if staggered and owner is not None and owner.size not in (1, phi.size):
        raise ValueError("precompute_grid_point_phi_prims: owner shape differs from phi")
    _lib.call("flub_phi_prims", phi.ptr, own, *phi.shape, *(float(s) for s in stagger), float(cp.h), n, xp, qp, gp,
              rp, None)


precompute_grid_point_phi_prims = Kernel("precompute_grid_point_phi_prims", _phi_prims,
                                         lambda phi, owner, x, q, g, st: phi.shape)


def _phi_to_fraction(phi, frac, dim):
    """geometry.py:264-290: frac (a MACGrid) holds face-centre phi on entry."""
    uf = cp.get_uf(frac, dim).f
    nx, ny, nz = (s - 1 for s in phi.shape)
    if uf.shape != cp.field_shape(dim, (nx, ny, nz)):
        raise ValueError("phi_to_fraction: face array does not match the corner array")
    _lib.call("flub_phi_to_fraction", phi.ptr, uf.ptr, nx, ny, nz, int(dim), None)


phi_to_fraction = Kernel("phi_to_fraction", _phi_to_fraction,
                         lambda phi, frac, dim: cp.get_uf(frac, dim).f.shape)


def _host3(a, n, what):
    v = np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1))
    if v.size != n:
        raise ValueError(f"update_argmin: {what} must hold {n} floats")
    return v


def _update_argmin(phi, owner, levelset, body_id, x, q, h_levelset, stagger, ijk, h=None):
    """geometry.py:365-388 (owner written only for a staggered array of phi's
    shape, as precompute_grid_point_phi_prims does)."""
    ls = levelset if isinstance(levelset, wp_dev.DeviceArray) else wp_dev.from_numpy(
        np.ascontiguousarray(levelset, np.float32))
    if len(ls.shape) != 3:
        raise ValueError("update_argmin: the level set must be a 3D array")
    xs, qs, st, ij = (_host3(x, 3, "x"), _host3(q, 4, "q"), _host3(stagger, 3, "stagger"), _host3(ijk, 3, "ijk"))
    staggered = any(float(v) != 0.0 for v in st)
    own = owner.ptr if (staggered and owner is not None and owner.size == phi.size) else None
    _lib.call("flub_update_argmin", phi.ptr, own, *phi.shape, ls.ptr, *ls.shape, int(body_id), xs.ctypes.data,
              qs.ctypes.data, float(h_levelset), st.ctypes.data, ij.ctypes.data,
              float(cp.h if h is None else h), None)


update_argmin = Kernel("update_argmin", _update_argmin,
                       lambda phi, owner, ls, i, x, q, hl, st, ijk, h=None: phi.shape)


class PrecomputeFaceFraction:
    """geometry.py:401-462: primitive bodies (boxes, spheres, "x") and bodies
    given as cached level sets.  ``vertices`` / ``edges`` are kept as the
    reference keeps them; no part of the projection path reads them."""

    def __init__(self, use_mesh=False, vertices=(), edges=(), h=None):
        self.vertices = np.array(vertices, dtype=np.float32)
        self.edges = np.array(edges, dtype=int)
        self.use_mesh = use_mesh
        self.h = h   # grid spacing; None: the module grid's h (const_param.h)

    def _h(self):
        return float(cp.h if self.h is None else self.h)

    def tag_dirichlet(self, tag, phi):
        _tag_dirichlet(tag, phi)

    def tag_neumann(self, tag, phi):
        _tag_neumann(tag, phi)

    def precompute_phi_with_cached_levelset(self, phi, rc, owner=None, stagger=(0.0, 0.0, 0.0)):
        """geometry.py:417-433: the primitive bodies, then each cached level
        set in the dict's order (update_argmin at the body's current pose)."""
        n, xp, qp, gp, rp = _bodies(rc.x_pair.cur, rc.theta_pair.cur, rc.rigid_geometry)
        staggered = any(float(s) != 0.0 for s in stagger)
        own = owner.ptr if (staggered and owner is not None and owner.size == phi.size) else None
        _lib.call("flub_phi_prims", phi.ptr, own, *phi.shape, *(float(s) for s in stagger), self._h(), n, xp, qp, gp,
                  rp, None)
        cached = getattr(rc, "cached_levelset", None) or {}
        if cached:
            xs = rc.x_pair.cur.numpy().reshape(-1, 3)
            qs = rc.theta_pair.cur.numpy().reshape(-1, 4)
            for i, (levelset, ijk, h_levelset) in cached.items():
                _update_argmin(phi, owner, levelset, i, xs[i], qs[i], h_levelset, stagger, ijk, self._h())

    def precompute_fij_with_cache(self, frac, vel_owner, rc):
        """geometry.py:438-462: fractions and owners of all faces.  With
        primitive bodies only, one fused pass per axis, bitwise the reference's
        staged corner/centre/fraction launches; with cached level sets, those
        staged launches (corner distances, face-centre distances and owners,
        phi_to_fraction per axis)."""
        fu, fv, fw = frac.u.f, frac.v.f, frac.w.f
        nx, ny, nz = fv.shape[0], fu.shape[1], fu.shape[2]
        for d, a in ((1, fu), (2, fv), (3, fw)):
            if a.shape != cp.field_shape(d, (nx, ny, nz)) or cp.get_uf(vel_owner, d).f.shape != a.shape:
                raise ValueError("precompute_fij_with_cache: face array shapes differ from the grid")
        if getattr(rc, "cached_levelset", None):
            corner = wp_dev.empty((nx + 1, ny + 1, nz + 1), float)
            self.precompute_phi_with_cached_levelset(corner, rc)
            for d, a, st in ((1, fu, (0.0, 0.5, 0.5)), (2, fv, (0.5, 0.0, 0.5)), (3, fw, (0.5, 0.5, 0.0))):
                self.precompute_phi_with_cached_levelset(a, rc, cp.get_uf(vel_owner, d).f, stagger=st)
            for d, a in ((1, fu), (2, fv), (3, fw)):
                _lib.call("flub_phi_to_fraction", corner.ptr, a.ptr, nx, ny, nz, d, None)
            return
        n, xp, qp, gp, rp = _bodies(rc.x_pair.cur, rc.theta_pair.cur, rc.rigid_geometry)
        _lib.call("flub_face_fractions", fu.ptr, fv.ptr, fw.ptr, vel_owner.u.f.ptr, vel_owner.v.f.ptr,
                  vel_owner.w.f.ptr, nx, ny, nz, self._h(), n, xp, qp, gp, rp, None)
