"""GPU parity of the solid-geometry kernels against the CPU oracle (through the
C-ABI): primitive distances and owners, face fractions (staged and fused),
rigid face velocities and the fraction-weighted divergence, all bitwise.
"""
import numpy as np
import pytest

from flub_amd import const_param as cp
from flub_amd import device as wp
from flub_amd.const_param import Field, make_mac_grid
from flub_amd.divergence import divergence_fractional_3d
from flub_amd.geometry import (GEOM_CIRCLE, GEOM_MESH, GEOM_NONE, GEOM_QUAD, GEOM_RECT, GEOM_X, GeometryType3D,
                               PrecomputeFaceFraction, phi_to_fraction, precompute_grid_point_phi_prims)
from flub_amd.launch import launch
from flub_amd.rigid_coupling import RigidBodies

pytestmark = pytest.mark.gpu

STAGGERS = {0: (0.5, 0.5, 0.5), 1: (0.0, 0.5, 0.5), 2: (0.5, 0.0, 0.5), 3: (0.5, 0.5, 0.0)}


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.dtype == np.float32:
        a, b = a.view(np.uint32), np.ascontiguousarray(b, np.float32).view(np.uint32)
    return a.shape == b.shape and np.array_equal(a, b)


def random_bodies(seed, n, kinds=(GEOM_QUAD, GEOM_CIRCLE, GEOM_RECT, GEOM_X)):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.2, 0.8, (n, 3)).astype(np.float32)
    q = rng.normal(size=(n, 4))
    q = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
    kinds = np.array([kinds[i % len(kinds)] for i in range(n)], np.int32)
    rad = rng.uniform(0.04, 0.15, (n, 3)).astype(np.float32)
    v = rng.normal(0, 0.3, (n, 3)).astype(np.float32)
    om = rng.normal(0, 0.3, (n, 3)).astype(np.float32)
    geoms = [GeometryType3D(int(k), tuple(float(r) for r in rr)) for k, rr in zip(kinds, rad)]
    return RigidBodies(x, q, geoms, v, om), (x, q, kinds, rad, v, om)


def oracle_bodies(oracle, arrs):
    x, q, kinds, rad, v, om = arrs
    return oracle.Rigids(x, q, kinds, rad, v, om)


@pytest.fixture
def grid():
    saved = cp.shape

    def set_(shape):
        cp.set_grid(*shape)
    yield set_
    cp.set_grid(*saved)


@pytest.mark.parametrize("shape,nb,seed", [((32, 32, 32), 3, 0), ((48, 40, 24), 6, 1), ((64, 64, 64), 1, 2)])
def test_phi_prims_and_owner_bitexact(shape, nb, seed, grid, oracle):
    grid(shape)
    rb, arrs = random_bodies(seed, nb)
    ob = oracle_bodies(oracle, arrs)
    nx, ny, nz = shape
    for d, st in STAGGERS.items():
        fs = cp.field_shape(d, shape)
        phi = wp.empty(fs, float)
        own = wp.empty(fs, int) if d else None
        launch(precompute_grid_point_phi_prims, dim=fs, inputs=[phi, own, rb.x_pair.cur, rb.theta_pair.cur,
                                                                 rb.rigid_geometry, st])
        ref_phi, ref_own = oracle.phi_prims(fs, st, cp.h, ob, owner=bool(d))
        assert bits_equal(phi.numpy(), ref_phi), f"phi differs on array {d}"
        if d:
            assert np.array_equal(own.numpy(), ref_own), f"owner differs on array {d}"
    # the corner array
    cs = (nx + 1, ny + 1, nz + 1)
    phi = wp.empty(cs, float)
    launch(precompute_grid_point_phi_prims, dim=cs, inputs=[phi, None, rb.x_pair.cur, rb.theta_pair.cur,
                                                             rb.rigid_geometry, (0.0, 0.0, 0.0)])
    assert bits_equal(phi.numpy(), oracle.phi_prims(cs, (0.0, 0.0, 0.0), cp.h, ob, owner=False)[0])


@pytest.mark.parametrize("shape,nb,seed", [((32, 32, 32), 3, 3), ((40, 24, 56), 5, 4), ((96, 96, 96), 2, 5)])
def test_face_fractions_staged_and_fused_bitexact(shape, nb, seed, grid, oracle):
    grid(shape)
    rb, arrs = random_bodies(seed, nb)
    (rfu, rfv, rfw), (rou, rov, row) = oracle.face_fractions(shape, cp.h, oracle_bodies(oracle, arrs))
    # fused (PrecomputeFaceFraction.precompute_fij_with_cache)
    frac = make_mac_grid(float, shape)
    owner = make_mac_grid(int, shape)
    PrecomputeFaceFraction().precompute_fij_with_cache(frac, owner, rb)
    for got, want in ((frac.u, rfu), (frac.v, rfv), (frac.w, rfw)):
        assert bits_equal(got.f.numpy(), want)
    for got, want in ((owner.u, rou), (owner.v, rov), (owner.w, row)):
        assert np.array_equal(got.f.numpy(), want)
    assert ((rfu >= 0) & (rfu <= 1)).all() and (rfu < 1).any() and (rfu == 1).any()
    # staged: corner phi, face-centre phi into the fraction arrays, phi_to_fraction
    nx, ny, nz = shape
    corner = wp.empty((nx + 1, ny + 1, nz + 1), float)
    pff = PrecomputeFaceFraction()
    pff.precompute_phi_with_cached_levelset(corner, rb)
    staged = make_mac_grid(float, shape)
    for d in (1, 2, 3):
        pff.precompute_phi_with_cached_levelset(cp.get_uf(staged, d).f, rb, cp.get_uf(owner, d).f, STAGGERS[d])
        launch(phi_to_fraction, dim=cp.field_shape(d, shape), inputs=[corner, staged, d])
    for got, want in ((staged.u, rfu), (staged.v, rfv), (staged.w, rfw)):
        assert bits_equal(got.f.numpy(), want)


def test_face_fractions_no_bodies_and_special_kinds(grid, oracle):
    shape = (16, 16, 16)
    grid(shape)
    rb = RigidBodies(np.zeros((0, 3)), np.zeros((0, 4)), [])
    frac = make_mac_grid(float, shape)
    owner = make_mac_grid(int, shape)
    PrecomputeFaceFraction().precompute_fij_with_cache(frac, owner, rb)
    assert (frac.u.f.numpy() == 1.0).all() and (owner.w.f.numpy() == -1).all()
    # "none" and mesh kinds contribute their fixed distances (1.0 / 2.0)
    rb, arrs = random_bodies(7, 4, kinds=(GEOM_NONE, GEOM_MESH, GEOM_QUAD, GEOM_CIRCLE))
    (rfu, rfv, rfw), _ = oracle.face_fractions(shape, cp.h, oracle_bodies(oracle, arrs))
    PrecomputeFaceFraction().precompute_fij_with_cache(frac, owner, rb)
    assert bits_equal(frac.v.f.numpy(), rfv)


@pytest.mark.parametrize("shape,seed", [((32, 32, 32), 8), ((64, 48, 40), 9)])
def test_solid_vel_bitexact(shape, seed, grid, oracle):
    grid(shape)
    rb, arrs = random_bodies(seed, 3)
    ob = oracle_bodies(oracle, arrs)
    frac = make_mac_grid(float, shape)
    owner = make_mac_grid(int, shape)
    PrecomputeFaceFraction().precompute_fij_with_cache(frac, owner, rb)
    vel = make_mac_grid(float, shape)
    rb.precompute_solid_vel_3d(rb.x_pair.cur, rb.theta_pair.cur, rb.v, rb.omega, rb.rigid_geometry, vel, owner)
    for d in (1, 2, 3):
        ow = cp.get_uf(owner, d).f.numpy()
        want = oracle.solid_vel(ow, d, cp.h0, ob)
        got = cp.get_uf(vel, d).f.numpy()
        assert bits_equal(got, want)
        assert (got != 0).any()


@pytest.mark.parametrize("shape,seed", [((32, 32, 32), 10), ((72, 40, 56), 11), ((128, 128, 128), 12)])
def test_divergence_fractional_bitexact(shape, seed, oracle):
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    fs = [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)]
    u, v, w = (rng.normal(0, 1, s).astype(np.float32) for s in fs)
    fu, fv, fw = (np.where(rng.random(s) < 0.7, 1.0, rng.random(s)).astype(np.float32) for s in fs)
    ur, vr, wr = (rng.normal(0, 0.3, s).astype(np.float32) for s in fs)
    dev = [wp.from_numpy(a) for a in (u, v, w)]
    div = wp.empty(shape, float)
    fr = [wp.from_numpy(a) for a in (fu, fv, fw)]
    rg = [wp.from_numpy(a) for a in (ur, vr, wr)]
    launch(divergence_fractional_3d, dim=shape, inputs=[*dev, div, *fr, *rg])
    assert bits_equal(div.numpy(), oracle.divergence_fractional(u, v, w, fu, fv, fw, ur, vr, wr))


def test_divergence_fractional_rejects_bad_shapes():
    a = wp.empty((9, 8, 8), float)
    div = wp.empty((8, 8, 8), float)
    with pytest.raises(ValueError):
        launch(divergence_fractional_3d, dim=(8, 8, 8), inputs=[a, a, a, div, a, a, a, a, a, a])


def _box_levelset(lshape, h_ls, half):
    """The signed distance of a box (half-extents ``half``) centred on the body
    origin, sampled on a grid of its own: index a sits at body coordinate
    (a + ijk) h_ls with ijk = -(l - 1) / 2 -- the form of a mesh body's cached
    level set (geometry.py:427-433)."""
    ijk = -(np.array(lshape, np.float64) - 1.0) / 2.0
    ax = [(np.arange(n) + o) * h_ls for n, o in zip(lshape, ijk)]
    p = np.stack(np.meshgrid(*ax, indexing="ij"), -1)
    m = np.abs(p) - np.array(half)
    out = np.where(m.max(-1) < 0, m.max(-1), np.linalg.norm(np.maximum(m, 0.0), axis=-1))
    return out.astype(np.float32), ijk.astype(np.float32)


def _mesh_body_scene(seed, nprim=3):
    """nprim primitive bodies plus one mesh body (kind 5, the last) whose
    distance comes from a cached level set."""
    rb, arrs = random_bodies(seed, nprim + 1, kinds=(GEOM_QUAD, GEOM_CIRCLE, GEOM_X, GEOM_MESH))
    x, q, kinds, rad, v, om = arrs
    kinds[-1] = GEOM_MESH
    geoms = [GeometryType3D(int(k), tuple(float(r) for r in rr)) for k, rr in zip(kinds, rad)]
    rb = RigidBodies(x, q, geoms, v, om)
    ls, ijk = _box_levelset((26, 22, 30), 0.02, (0.16, 0.12, 0.2))
    rb.cached_levelset = {nprim: (wp.from_numpy(ls), ijk, 0.02)}
    return rb, (x, q, kinds, rad, v, om), ls, ijk


@pytest.mark.parametrize("shape,seed", [((32, 32, 32), 21), ((40, 24, 56), 22)])
def test_update_argmin_bitexact(shape, seed, grid, oracle):
    """geometry.py:365-388: a cached level set resampled into distance arrays
    of every stagger (owners on the staggered ones), bitwise the oracle; the
    level set covers part of the grid, so clamped samples are included."""
    from flub_amd.geometry import update_argmin
    grid(shape)
    rb, arrs, ls, ijk = _mesh_body_scene(seed)
    ob = oracle_bodies(oracle, arrs)
    x, q = arrs[0][-1], arrs[1][-1]
    changed = 0
    for d, st in STAGGERS.items():
        fs = cp.field_shape(d, shape)
        phi0, own0 = oracle.phi_prims(fs, st, cp.h, ob, owner=True)
        phi, own = wp.from_numpy(phi0), wp.from_numpy(own0)
        launch(update_argmin, dim=fs, inputs=[phi, own, wp.from_numpy(ls), 3, x, q, 0.02, st, ijk, cp.h])
        want_phi, want_own = oracle.update_argmin(phi0, own0, ls, 3, x, q, 0.02, st, ijk, cp.h)
        assert bits_equal(phi.numpy(), want_phi), f"phi differs on array {d}"
        assert np.array_equal(own.numpy(), want_own), f"owner differs on array {d}"
        changed += int((want_phi != phi0).sum())
        if d:
            assert (want_own == 3).any()
    assert changed > 100


@pytest.mark.parametrize("shape,seed", [((32, 32, 32), 23), ((48, 40, 24), 24)])
def test_face_fractions_cached_levelset_bitexact(shape, seed, grid, oracle):
    """precompute_fij_with_cache with a cached mesh level set (the staged
    corner / face-centre / phi_to_fraction launches of geometry.py:438-462)
    against the oracle's same sequence, fractions and owners bitwise."""
    grid(shape)
    rb, arrs, ls, ijk = _mesh_body_scene(seed)
    ob = oracle_bodies(oracle, arrs)
    x, q = arrs[0][-1], arrs[1][-1]
    nx, ny, nz = shape
    cs = (nx + 1, ny + 1, nz + 1)
    corner, _ = oracle.phi_prims(cs, (0.0, 0.0, 0.0), cp.h, ob, owner=False)
    corner, _ = oracle.update_argmin(corner, None, ls, 3, x, q, 0.02, (0.0, 0.0, 0.0), ijk, cp.h)
    frac = make_mac_grid(float, shape)
    owner = make_mac_grid(int, shape)
    PrecomputeFaceFraction().precompute_fij_with_cache(frac, owner, rb)
    plain = make_mac_grid(float, shape)
    rb_plain = RigidBodies(*arrs[:2], [GeometryType3D(int(k), tuple(float(r) for r in rr))
                                       for k, rr in zip(arrs[2], arrs[3])])
    PrecomputeFaceFraction().precompute_fij_with_cache(plain, make_mac_grid(int, shape), rb_plain)
    for d in (1, 2, 3):
        fs = cp.field_shape(d, shape)
        cen, own = oracle.phi_prims(fs, STAGGERS[d], cp.h, ob, owner=True)
        cen, own = oracle.update_argmin(cen, own, ls, 3, x, q, 0.02, STAGGERS[d], ijk, cp.h)
        want = oracle.phi_to_fraction(corner, cen, d)
        got = cp.get_uf(frac, d).f.numpy()
        assert bits_equal(got, want), f"fractions differ on axis {d}"
        assert np.array_equal(cp.get_uf(owner, d).f.numpy(), own), f"owners differ on axis {d}"
        assert (own == 3).any() and not bits_equal(got, cp.get_uf(plain, d).f.numpy())
