"""Pins of the CPU oracle's solid-geometry routines (oracle/oracle_geometry.c)
against closed forms and independent libraries, never against a retyped copy
of the same formula:

* box / sphere / "x" distances (geometry.py:17-32,82-113) against the textbook
  box SDF of points rotated by scipy's Rotation (which also pins the
  quaternion rotation and product conventions);
* the argmin over bodies (geometry.py:148-169) against brute force, including
  its "closer than h0, else 2.0" rule;
* fraction_quad (utils/util.py:168-228) against the exact area of the unit
  square above a plane (Sutherland-Hodgman clipping), which the 4-triangle
  rule reproduces exactly for linear level sets;
* phi_to_fraction (geometry.py:264-290) per axis against that exact area on
  every face of a grid cut by a plane;
* update_argmin (geometry.py:365-388, cached mesh level sets) against a level
  set that is a cubic polynomial per axis, which the tricubic sampling
  reproduces exactly away from the clamped border, evaluated in FP64 at the
  point moved into the body frame by scipy's Rotation;
* precompute_solid_vel (rigid_coupling.py:51-70) against v + omega x (p - x);
* divergence_fractional_3d (divergence.py:5-41) against the plain flux
  divergence (all fractions 1), the rigid flux divergence (all fractions 0),
  u == u_rigid, and linearity.
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import oracle as O

F32 = np.float32


def _box_sdf(p, r):
    m = np.abs(p) - r
    mx = m.max(axis=-1)
    out = np.linalg.norm(np.maximum(m, 0.0), axis=-1)
    return np.where(mx < 0.0, mx, out)


def _rand_quat(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q.astype(F32)


def test_box_sdf_rotated_matches_scipy():
    rng = np.random.default_rng(0)
    qs = _rand_quat(rng, 20)
    for q in qs:
        r = rng.uniform(0.05, 0.3, 3).astype(F32)
        pts = rng.uniform(-0.5, 0.5, (50, 3)).astype(F32)
        rot = Rotation.from_quat(q.astype(np.float64))  # scalar-last (x, y, z, w)
        want = _box_sdf(rot.inv().apply(pts.astype(np.float64)), r.astype(np.float64))
        got = np.array([O.sdf_cube(p, r, q) for p in pts])
        np.testing.assert_allclose(got, want, atol=2e-6)


def test_sphere_and_none_and_mesh_prims():
    rng = np.random.default_rng(1)
    rb = O.Rigids(x=[[0.3, 0.4, 0.5], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]],
                  q=[[0, 0, 0, 1]] * 3, gtype=[2, 0, 5], radius=[[0.2, 0, 0], [0.1] * 3, [0.1] * 3])
    for p in rng.uniform(0, 1, (30, 3)).astype(F32):
        d = np.linalg.norm(p.astype(np.float64) - [0.3, 0.4, 0.5]) - 0.2
        assert abs(O.sdf_atom_prims(p, rb, 0) - d) < 2e-6
        assert O.sdf_atom_prims(p, rb, 1) == 1.0   # "none"
        assert O.sdf_atom_prims(p, rb, 2) == 2.0   # mesh: resolved from level sets elsewhere


def test_x_shape_is_union_of_box_and_its_quarter_turn():
    rng = np.random.default_rng(2)
    rz = Rotation.from_euler("z", 90, degrees=True)
    for q in _rand_quat(rng, 10):
        r = np.array([0.25, 0.04, 0.06], F32)
        rb = O.Rigids(x=[0.5, 0.5, 0.5], q=q, gtype=[4], radius=r)
        rq = Rotation.from_quat(q.astype(np.float64))
        r2 = rz * rq * rz.inv()
        pts = rng.uniform(0.2, 0.8, (40, 3)).astype(F32)
        rel = pts.astype(np.float64) - 0.5
        want = np.minimum(_box_sdf(rq.inv().apply(rel), r), _box_sdf(r2.inv().apply(rel), r))
        got = np.array([O.sdf_atom_prims(p, rb, 0) for p in pts])
        np.testing.assert_allclose(got, want, atol=3e-6)


def test_argmin_rule_and_owner_brute_force():
    rng = np.random.default_rng(3)
    n = 8
    h = 1.0 / 16
    x = rng.uniform(0.2, 0.8, (n, 3)).astype(F32)
    rad = rng.uniform(0.03, 0.12, (n, 3)).astype(F32)
    gt = rng.choice([1, 2, 3], n).astype(np.int32)
    q = _rand_quat(rng, n)
    rb = O.Rigids(x, q, gt, rad)
    shape = (17, 16, 16)
    phi, own = O.phi_prims(shape, (0.0, 0.5, 0.5), h, rb)
    for (i, j, k) in [tuple(rng.integers(0, s) for s in shape) for _ in range(300)]:
        p = np.array([F32(i) * F32(h) + F32(0.0) * F32(h), F32(j) * F32(h) + F32(0.5) * F32(h),
                      F32(k) * F32(h) + F32(0.5) * F32(h)], F32)
        d = np.array([O.sdf_atom_prims(p, rb, b) for b in range(n)])
        best = np.argmin(d)
        if d[best] < F32(h):
            assert phi[i, j, k] == F32(d[best]) and own[i, j, k] == best
        else:
            assert phi[i, j, k] == 2.0 and own[i, j, k] == -1
    # centred arrays carry no owner
    _, none = O.phi_prims((4, 4, 4), (0.5, 0.5, 0.5), h, rb, owner=False)
    assert none is None


def _clip_area(vals_fn):
    """Exact area of {f > 0} in the unit square for linear f (polygon clipping)."""
    poly = [(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)]
    out = []
    for a, b in zip(poly, poly[1:] + poly[:1]):
        fa, fb = vals_fn(*a), vals_fn(*b)
        if fa > 0:
            out.append(a)
        if (fa > 0) != (fb > 0):
            t = fa / (fa - fb)
            out.append((a[0] + t * (b[0] - a[0]), a[1] + t * (b[1] - a[1])))
    if len(out) < 3:
        return 0.0
    xs, ys = np.array(out).T
    return 0.5 * abs(np.dot(xs, np.roll(ys, -1)) - np.dot(ys, np.roll(xs, -1)))


def _edge_crossings(vals):
    """Positive-side fractions of the crossing edges of the 4-triangle fan."""
    ya, yb, yc, yd, ye = vals
    edges = [(ya, yb), (yb, yc), (yc, yd), (yd, ya), (ya, ye), (yb, ye), (yc, ye), (yd, ye)]
    out = []
    for p, q in edges:
        if (p > 0) != (q > 0):
            pos, neg = (p, q) if p > 0 else (q, p)
            out.append(pos / (pos - neg))
    return out


def test_fraction_quad_exact_for_planes():
    rng = np.random.default_rng(4)
    checked = 0
    for _ in range(4000):
        a, b = rng.normal(size=2)
        c = rng.uniform(-1.5, 1.5)
        f = lambda x, y: a * (x - 0.5) + b * (y - 0.5) + c * 0.5
        vals = [f(0, 0), f(1, 0), f(1, 1), f(0, 1), f(0.5, 0.5)]
        if min(abs(v) for v in vals) < 1e-3:
            continue
        cr = _edge_crossings(vals)
        area = _clip_area(f)
        if any(t < 0.11 for t in cr) or 0.0 < area < 0.11:
            continue
        got = O.fraction_quad(*[F32(v) for v in vals])
        assert abs(got - area) < 2e-5, (vals, got, area)
        checked += 1
    assert checked > 1500


def test_fraction_cutoffs():
    # an edge with less than 10% on the positive side counts as fully solid
    assert O.fraction_edge(0.05, -0.95) == 0.0
    assert O.fraction_edge(-0.95, 0.05) == 0.0
    assert abs(O.fraction_edge(0.2, -0.8) - 0.2) < 1e-7
    # a face with under 10% open area is closed; all-positive is open
    assert O.fraction_quad(1, 1, 1, 1, 1) == 1.0
    assert O.fraction_quad(-1, -1, -1, -1, -1) == 0.0
    assert O.fraction_quad(0.3, -1, -1, -1, -0.5) == 0.0
    # sdf == 0 counts as solid at every vertex
    assert O.fraction_quad(0, 0, 0, 0, 0) == 0.0


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_phi_to_fraction_plane_per_axis(dim):
    nx, ny, nz = 7, 6, 5
    h = 1.0 / 8
    nrm = np.array([0.6, -0.5, 0.62])
    off = 0.05

    def plane(x, y, z):
        return nrm[0] * x + nrm[1] * y + nrm[2] * z - off

    ii, jj, kk = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    corner = plane(ii * h, jj * h, kk * h).astype(F32)
    fs = (nx + (dim == 1), ny + (dim == 2), nz + (dim == 3))
    st = O.STAGGER[dim]
    fi, fj, fk = np.meshgrid(*(np.arange(s) for s in fs), indexing="ij")
    centre = plane((fi + st[0]) * h, (fj + st[1]) * h, (fk + st[2]) * h).astype(F32)
    frac = O.phi_to_fraction(corner, centre, dim)
    ax = [a for a in range(3) if a != dim - 1]   # the face spans these two axes
    checked = 0
    for idx in np.ndindex(*fs):
        base = np.array(idx, np.float64) * h

        def f(s, t):
            p = base.copy()
            p[ax[0]] += s * h
            p[ax[1]] += t * h
            return plane(*p)

        vals = [f(0, 0), f(1, 0), f(1, 1), f(0, 1), f(0.5, 0.5)]
        if min(abs(v) for v in vals) < 1e-4:
            continue
        area = _clip_area(f)
        if any(t < 0.11 for t in _edge_crossings(vals)) or 0.0 < area < 0.11:
            continue
        assert abs(frac[idx] - area) < 5e-5, (dim, idx, frac[idx], area)
        checked += 1
    assert checked > 50


def test_solid_vel_rigid_motion():
    rng = np.random.default_rng(5)
    n = 16
    h0 = 1.0 / n
    rb = O.Rigids(x=[[0.4, 0.5, 0.6], [0.7, 0.2, 0.3]], q=[[0, 0, 0, 1]] * 2, gtype=[1, 1],
                  radius=[[0.1] * 3] * 2, v=rng.normal(0, 0.3, (2, 3)), omega=rng.normal(0, 0.3, (2, 3)))
    for d, fs in ((1, (n + 1, n, n)), (2, (n, n + 1, n)), (3, (n, n, n + 1))):
        owner = rng.integers(-1, 2, fs).astype(np.int32)
        got = O.solid_vel(owner, d, h0, rb)
        st = np.array(O.STAGGER[d])
        for idx in np.ndindex(*fs):
            o = owner[idx]
            t = idx[0] * h0
            if o == -1 or not (0.01 * h0 < t < 1 - 0.01 * h0):
                assert got[idx] == 0.0
                continue
            p = (np.array(idx) + st) * h0
            want = rb.v[o].astype(np.float64) + np.cross(rb.omega[o].astype(np.float64), p - rb.x[o])
            assert abs(got[idx] - want[d - 1]) < 2e-6


def _faces(rng, n, lo=-1.0, hi=1.0):
    return (rng.uniform(lo, hi, (n + 1, n, n)).astype(F32), rng.uniform(lo, hi, (n, n + 1, n)).astype(F32),
            rng.uniform(lo, hi, (n, n, n + 1)).astype(F32))


def test_divergence_fractional_special_cases():
    rng = np.random.default_rng(6)
    n = 12
    u, v, w = _faces(rng, n)
    ur, vr, wr = _faces(rng, n)
    ones = tuple(np.ones_like(a) for a in (u, v, w))
    zeros = tuple(np.zeros_like(a) for a in (u, v, w))
    fluid = np.ones((n, n, n), np.int32)
    sl = (slice(1, -1),) * 3   # divergence_tag zeroes the wall faces
    # all fractions 1: the plain flux divergence (the already pinned divergence_tag)
    d1 = O.divergence_fractional(u, v, w, *ones, ur, vr, wr)
    assert np.array_equal(d1[sl], O.divergence_tag(u, v, w, ur, vr, wr, fluid)[sl])
    # all fractions 0: the same divergence of the rigid velocities
    d0 = O.divergence_fractional(u, v, w, *zeros, ur, vr, wr)
    assert np.array_equal(d0[sl], O.divergence_tag(ur, vr, wr, u, v, w, fluid)[sl])
    # any fractions with u == u_rigid: the plain divergence of u
    fr = _faces(rng, n, 0.0, 1.0)
    ds = O.divergence_fractional(u, v, w, *fr, u, v, w)
    np.testing.assert_allclose(ds[sl], d1[sl], atol=5e-6)
    # linearity in (u, u_rigid) at fixed fractions
    da = O.divergence_fractional(u, v, w, *fr, *zeros)
    db = O.divergence_fractional(*zeros, *fr, ur, vr, wr)
    dc = O.divergence_fractional(u, v, w, *fr, ur, vr, wr)
    np.testing.assert_allclose(dc, da + db, atol=5e-6)


def _cubic_levelset(a, b, c):
    """A tensor-product polynomial of degree 3 per axis in level-set index units."""
    return (0.05 + 0.004 * a - 0.003 * b + 0.002 * c + 1e-4 * a * b - 2e-5 * c * c
            + 3e-6 * a ** 3 - 1e-6 * b ** 2 * c + 1e-7 * a * b * c ** 2)


@pytest.mark.parametrize("stagger", [(0.0, 0.0, 0.0), (0.0, 0.5, 0.5), (0.5, 0.5, 0.0)])
def test_update_argmin_cached_levelset_polynomial(stagger):
    rng = np.random.default_rng(11)
    shape, h = (24, 20, 28), 1.0 / 24
    lshape, h_ls = (30, 30, 30), 0.025
    ijk = np.array([-14.0, -15.5, -13.0], F32)
    x = np.array([0.47, 0.41, 0.56], F32)
    q = _rand_quat(rng, 1)[0]
    a, b, c = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in lshape), indexing="ij")
    ls = _cubic_levelset(a, b, c).astype(F32)
    phi0 = rng.uniform(-0.05, 0.25, shape).astype(F32)
    own0 = rng.integers(-1, 3, shape).astype(np.int32)
    phi, own = O.update_argmin(phi0, own0, ls, 7, x, q, h_ls, stagger, ijk, h)

    i, j, k = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in shape), indexing="ij")
    pts = np.stack([i + stagger[0], j + stagger[1], k + stagger[2]], -1) * h - x.astype(np.float64)
    body = Rotation.from_quat(q.astype(np.float64)).inv().apply(pts.reshape(-1, 3)).reshape(pts.shape)
    s = body / h_ls - ijk.astype(np.float64)
    inside = np.all((s >= 1.0) & (s <= np.array(lshape) - 3.0), axis=-1)
    want = _cubic_levelset(s[..., 0], s[..., 1], s[..., 2])
    clear = inside & (np.abs(want - phi0) > 1e-4)
    assert clear.sum() > 500 and (clear & (want < phi0)).sum() > 100 and (clear & (want > phi0)).sum() > 100
    np.testing.assert_allclose(phi[clear], np.minimum(want, phi0)[clear], atol=2e-6)
    upd = clear & (want < phi0)
    keep = clear & (want > phi0)
    assert np.array_equal(phi[keep], phi0[keep])
    staggered = any(v != 0.0 for v in stagger)
    assert np.all(own[upd] == (7 if staggered else own0[upd]))
    assert np.array_equal(own[keep], own0[keep])
    # no owner array: the same distances
    phi2, _ = O.update_argmin(phi0, None, ls, 7, x, q, h_ls, stagger, ijk, h)
    assert np.array_equal(phi2.view(np.uint32), phi.view(np.uint32))
