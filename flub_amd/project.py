"""The projection harness: ``Example.divergence``/``call_mgpcg``/``project`` of
the reference (example_fluid.py:43-49,84-94,120-147; water_3d.py:58-68),
restricted to the hot path: from (face velocities, cell tags, rigid face
velocities) to (projected velocities, pressure).

``Projection.project()`` runs it through one C-ABI call (flub_project), which
fuses divergence/mask/scale and the pressure scale chain into the grid
kernels; ``Projection.project_stepwise()`` issues the reference's launch
sequence one kernel at a time through the drop-in modules.  Both produce the
same fields.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from . import device as wp
from .const_param import Field, IntField, MACGrid, field_shape
from .divergence import divergence_tag, mask_with_tag
from .launch import launch
from .multigrid.mgpcg import MGPCG, SMOOTHERS
from .subtract_gradient import subtract_gradient_tag
from .utils.util import scale_by_3d


def _field(arr, d):
    return Field(f=wp.from_numpy(np.asarray(arr, np.float32)), d=d)


# The tuned V-cycle (no reference counterpart): levels down to 4 cells on the
# shortest axis, the coarse-grid correction scaled by 1/2 (each restriction
# stores half the children's residual sum), at most 3 red-black iterations of
# pre- and post-smoothing on any level; same smoother, transfers and bottom
# schedule as the reference's MGPCG otherwise.
TUNED_CYCLE = {"coarse_scale": 0.5, "sweep_cap": 3}


def tuned_levels(shape):
    """Levels of the tuned cycle for a grid: halve until the shortest axis has
    4 cells, every restricted level even (as the V-cycle requires) and every
    level's z extent a multiple of 4 (as the 16-bit storage requires)."""
    shape = tuple(int(s) for s in shape)
    m = min(shape)
    levels = 1
    while (levels < 12 and m % 2 == 0 and m // 2 >= 4 and all((s >> (levels - 1)) % 2 == 0 for s in shape)
           and (shape[2] >> levels) % 4 == 0):
        m //= 2
        levels += 1
    return levels


class Projection:
    """Device state of one projection problem on an (nx, ny, nz) grid."""

    def __init__(self, shape, n_mg_levels=4, pre_and_post_smoothing=2, bottom_smoothing=6,
                 precision="fp32", h=None, coarse_scale=1.0, sweep_cap=0):
        self.shape = tuple(int(s) for s in shape)
        self.h = float(h) if h is not None else 1.0 / self.shape[0]
        self.mgpcg = MGPCG(3, self.shape, n_mg_levels, float, precision=precision)
        self.mgpcg.pre_and_post_smoothing = pre_and_post_smoothing
        self.mgpcg.bottom_smoothing = bottom_smoothing
        if coarse_scale != 1.0:
            self.mgpcg.coarse_scale = coarse_scale
        if sweep_cap:
            self.mgpcg.sweep_cap = sweep_cap
        g = self.shape
        self.vel0 = MACGrid(*(Field(f=wp.zeros(field_shape(d, g)), d=d) for d in (1, 2, 3)))
        self.vel_rigid = MACGrid(*(Field(f=wp.zeros(field_shape(d, g)), d=d) for d in (1, 2, 3)))
        self.tag = IntField(f=wp.zeros(g, int))
        self.div = Field(f=wp.zeros(g))
        self.p0 = Field(f=wp.zeros(g))
        self.iterations = None
        self.rTr = None

    # ---------------------------------------------------------------- inputs
    def load(self, scene):
        """Copy a scene dict (flub_amd.scenes) to the device."""
        self.set_velocity(scene["u"], scene["v"], scene["w"])
        self.set_rigid_velocity(scene["ur"], scene["vr"], scene["wr"])
        self.tag.f.assign(np.asarray(scene["tag"], np.int32))

    def set_velocity(self, u, v, w):
        self.vel0.u.f.assign(u)
        self.vel0.v.f.assign(v)
        self.vel0.w.f.assign(w)

    def set_rigid_velocity(self, u, v, w):
        self.vel_rigid.u.f.assign(u)
        self.vel_rigid.v.f.assign(v)
        self.vel_rigid.w.f.assign(w)

    def velocity(self):
        return self.vel0.u.f.numpy(), self.vel0.v.f.numpy(), self.vel0.w.f.numpy()

    # ---------------------------------------------------------------- fused
    def project(self, max_iters=-1, rel_tol=1e-12, abs_tol=0.0, smoother="phases"):
        it = ctypes.c_int(0)
        rtr = ctypes.c_double(0.0)
        v, r = self.vel0, self.vel_rigid
        _lib.call("flub_project", self.mgpcg.handle, v.u.f.ptr, v.v.f.ptr, v.w.f.ptr, r.u.f.ptr,
                  r.v.f.ptr, r.w.f.ptr, self.tag.f.ptr, float(self.h), int(max_iters), float(abs_tol),
                  float(rel_tol), SMOOTHERS[smoother], self.p0.f.ptr, self.div.f.ptr, ctypes.byref(it),
                  ctypes.byref(rtr), allow=(_lib.ENOTCONV,))
        self.iterations, self.rTr = it.value, rtr.value
        return self.iterations

    # ---------------------------------------------------------------- stepwise
    def divergence(self):
        """example_fluid.py:84-93 (without the rigid precompute and the print)."""
        shape = self.shape
        launch(divergence_tag, dim=shape, inputs=[self.vel0, self.div, self.vel_rigid, self.tag])
        launch(mask_with_tag, dim=shape, inputs=[self.div, self.tag])
        launch(scale_by_3d, dim=shape, inputs=[self.div, -1.0 / self.h])

    def call_mgpcg(self, pf, max_iters=-1, rel_tol=1e-12, abs_tol=0.0, smoother="phases", verbose=False):
        """water_3d.py:58-68."""
        mg = self.mgpcg
        mg.set_tags(self.tag.f)
        mg.init(self.div.f, -self.h)
        mg.mask_interior()
        self.iterations = mg.solve(max_iters, abs_tol=abs_tol, rel_tol=rel_tol, verbose=verbose,
                                   smoother=smoother)
        self.rTr = mg.last_rTr
        wp.copy(pf.f, mg.x)
        launch(scale_by_3d, dim=self.shape, inputs=[pf, self.h])

    def project_learned(self, solver, fluid, rel_tol=1e-12, abs_tol=0.0):
        """project() with the pressure solve done by a flub_amd.learned.fpcg.LearnedPCG
        (the U-Net preconditioner with MGPCG fallback) on this projection's
        MGPCG handle; fluid is a torch bool mask of the FLUID cells.  Returns
        the solver statistics."""
        self.divergence()
        return self._solve_learned_and_subtract(solver, fluid, rel_tol, abs_tol)

    # ---------------------------------------------------------- solid bodies
    def set_bodies(self, bodies):
        """Rigid bodies of the scene (dict of x, q, gtype, radius, v, omega
        arrays) for the face fractions and rigid face velocities."""
        from .const_param import make_mac_grid
        from .geometry import GeometryType3D, PrecomputeFaceFraction
        from .rigid_coupling import RigidBodies

        gt = np.asarray(bodies["gtype"], np.int32).reshape(-1)
        rad = np.asarray(bodies["radius"], np.float32).reshape(-1, 3)
        geoms = [GeometryType3D(int(k), tuple(float(v) for v in r)) for k, r in zip(gt, rad)]
        self.rc = RigidBodies(bodies["x"], bodies["q"], geoms, bodies.get("v"), bodies.get("omega"), h0=self.h)
        self.pff = PrecomputeFaceFraction(h=self.h)
        self.frac = make_mac_grid(float, self.shape)
        self.vel_owner = make_mac_grid(int, self.shape)

    def precompute_geometry(self):
        """The solid inputs of a projection: face fractions with their owning
        bodies (example_fluid.py:109, geometry.py:438-462) and the rigid face
        velocities (example_fluid.py:87, rigid_coupling.py:225-228)."""
        rc = self.rc
        self.pff.precompute_fij_with_cache(self.frac, self.vel_owner, rc)
        rc.precompute_solid_vel_3d(rc.x_pair.cur, rc.theta_pair.cur, rc.v, rc.omega, rc.rigid_geometry,
                                   self.vel_rigid, self.vel_owner)

    def divergence_fractional(self):
        """divergence() with the fraction-weighted fluxes of divergence.py:5-41
        (the launch example_fluid.py:88 names), then mask and scale."""
        from .divergence import divergence_fractional_3d

        v, f, r = self.vel0, self.frac, self.vel_rigid
        launch(divergence_fractional_3d, dim=self.shape,
               inputs=[v.u.f, v.v.f, v.w.f, self.div.f, f.u.f, f.v.f, f.w.f, r.u.f, r.v.f, r.w.f])
        launch(mask_with_tag, dim=self.shape, inputs=[self.div, self.tag])
        launch(scale_by_3d, dim=self.shape, inputs=[self.div, -1.0 / self.h])

    def project_fractional(self, solver=None, fluid=None, max_iters=-1, rel_tol=1e-12, abs_tol=0.0):
        """A projection of a scene with solid bodies: geometry, fraction-weighted
        divergence, the pressure solve (this handle's MGPCG, or a LearnedPCG
        `solver` with the torch FLUID mask `fluid`) and the gradient subtraction.
        Returns the iteration count, or the learned solver's statistics."""
        self.precompute_geometry()
        self.divergence_fractional()
        if solver is not None:
            return self._solve_learned_and_subtract(solver, fluid, rel_tol, abs_tol)
        self.p0.f.zero_()
        self.call_mgpcg(self.p0, max_iters, rel_tol, abs_tol)
        self._subtract_gradient()
        return self.iterations

    def _subtract_gradient(self):
        """example_fluid.py:134-145 after the solve: p0 / h, the three face
        updates, p0 * h."""
        launch(scale_by_3d, dim=self.shape, inputs=[self.p0, 1.0 / self.h])
        g = self.shape
        for d in (1, 2, 3):
            launch(subtract_gradient_tag, dim=field_shape(d, g), inputs=[self.vel0, self.vel_rigid, self.p0, self.tag, d])
        launch(scale_by_3d, dim=self.shape, inputs=[self.p0, self.h])

    def _solve_learned_and_subtract(self, solver, fluid, rel_tol, abs_tol):
        import torch

        mg = self.mgpcg
        mg.set_tags(self.tag.f)
        mg.init(self.div.f, -self.h)
        mg.mask_interior()
        b = torch.empty(self.shape, device="cuda", dtype=torch.float32)
        _lib.call("flub_memcpy_d2d", b.data_ptr(), mg.r[0].ptr, mg.r[0].nbytes, None)
        x, st = solver.solve(b, fluid, rel_tol=rel_tol, abs_tol=abs_tol)
        _lib.call("flub_memcpy_d2d", mg.x.ptr, x.data_ptr(), mg.x.nbytes, None)
        self.iterations, self.rTr = st["iterations"], st["true_rTr"]
        wp.copy(self.p0.f, mg.x)
        launch(scale_by_3d, dim=self.shape, inputs=[self.p0, self.h])
        # the gradient of the solver's FP64 x (p0 / h of the scale chain), as
        # the fused projection subtracts it
        x64 = solver.last_x64
        g = self.shape
        for d, uf, ur in ((1, self.vel0.u, self.vel_rigid.u), (2, self.vel0.v, self.vel_rigid.v),
                          (3, self.vel0.w, self.vel_rigid.w)):
            _lib.call("flub_subtract_gradient_tag_f64", uf.f.ptr, ur.f.ptr, x64.data_ptr(), self.tag.f.ptr, *g, d,
                      None)
        return st

    def project_stepwise(self, max_iters=-1, rel_tol=1e-12, abs_tol=0.0, smoother="phases", verbose=False):
        """example_fluid.py:120-145 launch by launch."""
        self.divergence()
        self.p0.f.zero_()
        self.call_mgpcg(self.p0, max_iters, rel_tol, abs_tol, smoother, verbose)
        self._subtract_gradient()
        return self.iterations
