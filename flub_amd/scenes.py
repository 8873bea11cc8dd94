"""Seeded scenes for the projection workloads (host-side NumPy input generation).

Unit cube, n^3 cells, h = 1/n, cell centre ((i+.5)h, (j+.5)h, (k+.5)h).  Each
scene draws from one ``numpy.random.default_rng(seed)`` stream in a fixed
order: geometry first, then the u, v, w face arrays in full, then masking.
A face is a "fluid face" when at least one of its two cells is FLUID; faces
on the domain walls are 0.  Level sets follow the reference's conventions:
water phi > 0 is AIR (geometry.py:335-344), rigid phi < 0 is SOLID
(geometry.py:390-397), and tags are produced from them in that order.

Scenes return a dict with ``phi`` (water level set), ``phi_rigid``, ``tag``
(int32), ``u, v, w`` and ``ur, vr, wr`` (rigid face velocities), ``h``.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32


def centres(n):
    return ((np.arange(n, dtype=np.float64) + 0.5) / n).astype(F32)


def _grid(n):
    c = centres(n)
    return c[:, None, None], c[None, :, None], c[None, None, :]


def tags_from_phi(phi, phi_rigid):
    """geometry.py:335-344 then :390-397 (NumPy twin of the HIP tag kernels)."""
    tag = np.where(phi > 0.0, 0, 1).astype(np.int32)
    tag[phi_rigid < 0.0] = 2
    return tag


def _sphere_phi(n, centres_xyz, radii, phi=None):
    """min over spheres of |c - x| - r, float32, accumulated sphere by sphere."""
    x, y, z = _grid(n)
    if phi is None:
        phi = np.full((n, n, n), np.inf, dtype=F32)
    for (cx, cy, cz), r in zip(centres_xyz, radii):
        d = np.sqrt((x - F32(cx)) ** 2 + (y - F32(cy)) ** 2 + (z - F32(cz)) ** 2) - F32(r)
        np.minimum(phi, d.astype(F32), out=phi)
    return phi


def _box_phi(n, xmax, ymax):
    """Dam-break block {x < xmax, y < ymax}: phi = max(x - xmax, y - ymax)."""
    x, y, _ = _grid(n)
    return np.broadcast_to(np.maximum(x - F32(xmax), y - F32(ymax)), (n, n, n)).astype(F32)


def fluid_face_masks(tag):
    """Boolean masks (u, v, w face shapes): at least one adjacent cell FLUID,
    wall faces excluded."""
    fl = tag == 1
    n = tag.shape
    mu = np.zeros((n[0] + 1, n[1], n[2]), bool)
    mv = np.zeros((n[0], n[1] + 1, n[2]), bool)
    mw = np.zeros((n[0], n[1], n[2] + 1), bool)
    mu[1:-1] = fl[1:] | fl[:-1]
    mv[:, 1:-1] = fl[:, 1:] | fl[:, :-1]
    mw[:, :, 1:-1] = fl[:, :, 1:] | fl[:, :, :-1]
    return mu, mv, mw


def solid_face_masks(tag):
    so = tag == 2
    n = tag.shape
    mu = np.zeros((n[0] + 1, n[1], n[2]), bool)
    mv = np.zeros((n[0], n[1] + 1, n[2]), bool)
    mw = np.zeros((n[0], n[1], n[2] + 1), bool)
    mu[1:-1] = so[1:] | so[:-1]
    mv[:, 1:-1] = so[:, 1:] | so[:, :-1]
    mw[:, :, 1:-1] = so[:, :, 1:] | so[:, :, :-1]
    return mu, mv, mw


def _faces_zero(n):
    return (np.zeros((n + 1, n, n), F32), np.zeros((n, n + 1, n), F32), np.zeros((n, n, n + 1), F32))


def _pack(n, phi, phi_rigid, u, v, w, ur, vr, wr, **extra):
    tag = tags_from_phi(phi, phi_rigid)
    d = dict(n=n, h=1.0 / n, phi=phi, phi_rigid=phi_rigid, tag=tag, u=u, v=v, w=w, ur=ur, vr=vr, wr=wr)
    d.update(extra)
    return d


def dam_break(n=64, seed=0, dt=1.0 / 60.0):
    """FLUID where x < 0.4 and y < 0.6, AIR elsewhere, walls closed; fluid faces
    U[-1, 1] plus -9.8*dt on v."""
    rng = np.random.default_rng(seed)
    phi = _box_phi(n, 0.4, 0.6)
    phi_rigid = np.ones((n, n, n), F32)
    u = rng.uniform(-1.0, 1.0, (n + 1, n, n)).astype(F32)
    v = rng.uniform(-1.0, 1.0, (n, n + 1, n)).astype(F32)
    w = rng.uniform(-1.0, 1.0, (n, n, n + 1)).astype(F32)
    v = (v + F32(-9.8 * dt)).astype(F32)
    tag = tags_from_phi(phi, phi_rigid)
    mu, mv, mw = fluid_face_masks(tag)
    u[~mu] = 0.0
    v[~mv] = 0.0
    w[~mw] = 0.0
    ur, vr, wr = _faces_zero(n)
    return _pack(n, phi, phi_rigid, u, v, w, ur, vr, wr)


def dam_break_spheres(n=128, seed=1, n_spheres=32, r_cells=(2.0, 8.0), sigma=0.5):
    """Dam-break block union n_spheres fluid spheres (radius U[r_cells] cells,
    centres U[0,1]^3); fluid faces N(0, sigma), all other faces 0."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, 1.0, (n_spheres, 3))
    r = rng.uniform(r_cells[0], r_cells[1], n_spheres) / n
    phi = _sphere_phi(n, c, r, _box_phi(n, 0.4, 0.6).copy())
    phi_rigid = np.ones((n, n, n), F32)
    u = rng.normal(0.0, sigma, (n + 1, n, n)).astype(F32)
    v = rng.normal(0.0, sigma, (n, n + 1, n)).astype(F32)
    w = rng.normal(0.0, sigma, (n, n, n + 1)).astype(F32)
    tag = tags_from_phi(phi, phi_rigid)
    mu, mv, mw = fluid_face_masks(tag)
    u[~mu] = 0.0
    v[~mv] = 0.0
    w[~mw] = 0.0
    ur, vr, wr = _faces_zero(n)
    return _pack(n, phi, phi_rigid, u, v, w, ur, vr, wr)


def flow_past_sphere(n=256, seed=2, radius=0.15, u_in=1.0, sigma=0.1):
    """Solid sphere (radius 0.15) at the centre; inflow u = u_in through a SOLID
    ghost layer i = 0 whose rigid face velocity on face i = 1 is u_in; outflow
    p = 0 through an AIR layer i = n-1; FLUID elsewhere.  u = u_in + N(0, sigma),
    v, w = N(0, sigma) on fluid faces; other faces 0 (solid faces take the rigid
    velocity inside the kernels)."""
    rng = np.random.default_rng(seed)
    x, y, z = _grid(n)
    phi = np.full((n, n, n), -1.0, F32)
    phi[n - 1] = 1.0                                   # AIR outflow layer
    dist = np.sqrt((x - F32(0.5)) ** 2 + (y - F32(0.5)) ** 2 + (z - F32(0.5)) ** 2)
    phi_rigid = np.broadcast_to((dist - F32(radius)).astype(F32), (n, n, n)).copy()
    phi_rigid[0] = -1.0                                # SOLID inflow ghost layer
    u = (F32(u_in) + rng.normal(0.0, sigma, (n + 1, n, n))).astype(F32)
    v = rng.normal(0.0, sigma, (n, n + 1, n)).astype(F32)
    w = rng.normal(0.0, sigma, (n, n, n + 1)).astype(F32)
    tag = tags_from_phi(phi, phi_rigid)
    mu, mv, mw = fluid_face_masks(tag)
    u[~mu] = 0.0
    v[~mv] = 0.0
    w[~mw] = 0.0
    ur, vr, wr = _faces_zero(n)
    ur[1] = F32(u_in)                                  # inflow through the ghost layer
    return _pack(n, phi, phi_rigid, u, v, w, ur, vr, wr)


def random_pool(n=128, seed=1000):
    """Training/evaluation scenes of the learned preconditioner: water level
    U[0.2, 0.7], 1-40 fluid spheres of radius U[3, 20] cells, 0-4 solid spheres
    of radius U[3, 20] cells; fluid faces N(0, 0.5)."""
    rng = np.random.default_rng(seed)
    level = rng.uniform(0.2, 0.7)
    nf = int(rng.integers(1, 41))
    ns = int(rng.integers(0, 5))
    cf = rng.uniform(0.0, 1.0, (nf, 3))
    rf = rng.uniform(3.0, 20.0, nf) / n
    cs = rng.uniform(0.0, 1.0, (ns, 3))
    rs = rng.uniform(3.0, 20.0, ns) / n
    _, y, _ = _grid(n)
    phi = np.broadcast_to((y - F32(level)).astype(F32), (n, n, n)).copy()
    phi = _sphere_phi(n, cf, rf, phi)
    phi_rigid = _sphere_phi(n, cs, rs) if ns else np.ones((n, n, n), F32)
    u = rng.normal(0.0, 0.5, (n + 1, n, n)).astype(F32)
    v = rng.normal(0.0, 0.5, (n, n + 1, n)).astype(F32)
    w = rng.normal(0.0, 0.5, (n, n, n + 1)).astype(F32)
    tag = tags_from_phi(phi, phi_rigid)
    mu, mv, mw = fluid_face_masks(tag)
    u[~mu] = 0.0
    v[~mv] = 0.0
    w[~mw] = 0.0
    ur, vr, wr = _faces_zero(n)
    return _pack(n, phi, phi_rigid, u, v, w, ur, vr, wr)


def _sphere_phi_local(n, centres_xyz, radii, phi):
    """_sphere_phi restricted to each sphere's bounding box (one cell of
    margin): phi is min(phi, |c - x| - r) there and unchanged elsewhere, so its
    sign (the tag) equals the full-grid min for spheres of any size."""
    c1 = centres(n)
    for (cx, cy, cz), r in zip(centres_xyz, radii):
        lo = [max(int(np.floor((c - r) * n)) - 1, 0) for c in (cx, cy, cz)]
        hi = [min(int(np.ceil((c + r) * n)) + 1, n) for c in (cx, cy, cz)]
        if any(a >= b for a, b in zip(lo, hi)):
            continue
        x = c1[lo[0]:hi[0], None, None]
        y = c1[None, lo[1]:hi[1], None]
        z = c1[None, None, lo[2]:hi[2]]
        d = np.sqrt((x - F32(cx)) ** 2 + (y - F32(cy)) ** 2 + (z - F32(cz)) ** 2) - F32(r)
        sub = phi[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
        np.minimum(sub, d.astype(F32), out=sub)
    return phi


def _random_quat(rng):
    q = rng.normal(size=4)
    return (q / np.linalg.norm(q)).astype(F32)


def _box_sdf_host(n, centre, q, half):
    """Signed distance of a box (half-extents `half`) at `centre` rotated by the
    unit quaternion q = (x, y, z, w), at the cell centres, in float64."""
    qx, qy, qz, qw = (float(a) for a in q)
    # rotation matrix of q; the body frame coordinates are R^T (p - c)
    R = np.array([[1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)],
                  [2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)],
                  [2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)]])
    c = (np.arange(n, dtype=np.float64) + 0.5) / n
    out = np.empty((n, n, n), F32)
    yy, zz = np.meshgrid(c - centre[1], c - centre[2], indexing="ij")
    for i in range(n):
        p = np.stack([np.full_like(yy, c[i] - centre[0]), yy, zz], axis=-1) @ R   # (R^T p) per point
        m = np.abs(p) - np.asarray(half, np.float64)
        mx = m.max(axis=-1)
        out[i] = np.where(mx < 0.0, mx, np.linalg.norm(np.maximum(m, 0.0), axis=-1))
    return out


def dam_break_droplets_box(n=512, seed=3, n_drops=200, r_cells=(2.0, 12.0), box=0.1, sigma_box=0.3, sigma=0.5):
    """The combined scene: the dam-break block (x < 0.4, y < 0.6) union n_drops
    fluid droplets (radius U[r_cells] cells, centres U[0,1]^3), and a solid box
    of side `box` at a random pose (centre U[0.3, 0.7]^3, uniformly random
    orientation) whose linear and angular velocities are N(0, sigma_box) per
    component.  Fluid faces N(0, sigma).  Besides the usual arrays the scene
    carries the body ('bodies': x, q, gtype, radius, v, omega) from which the
    face fractions and rigid face velocities are computed on the device."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, 1.0, (n_drops, 3))
    r = rng.uniform(r_cells[0], r_cells[1], n_drops) / n
    bc = rng.uniform(0.3, 0.7, 3).astype(F32)
    bq = _random_quat(rng)
    bv = rng.normal(0.0, sigma_box, 3).astype(F32)
    bw = rng.normal(0.0, sigma_box, 3).astype(F32)
    phi = _sphere_phi_local(n, c, r, _box_phi(n, 0.4, 0.6).copy())
    half = np.full(3, box / 2.0, F32)
    phi_rigid = _box_sdf_host(n, bc.astype(np.float64), bq, half)
    u = rng.normal(0.0, sigma, (n + 1, n, n)).astype(F32)
    v = rng.normal(0.0, sigma, (n, n + 1, n)).astype(F32)
    w = rng.normal(0.0, sigma, (n, n, n + 1)).astype(F32)
    tag = tags_from_phi(phi, phi_rigid)
    mu, mv, mw = fluid_face_masks(tag)
    u[~mu] = 0.0
    v[~mv] = 0.0
    w[~mw] = 0.0
    del mu, mv, mw
    ur, vr, wr = _faces_zero(n)
    bodies = dict(x=bc.reshape(1, 3), q=bq.reshape(1, 4), gtype=np.array([1], np.int32), radius=half.reshape(1, 3),
                  v=bv.reshape(1, 3), omega=bw.reshape(1, 3))
    return _pack(n, phi, phi_rigid, u, v, w, ur, vr, wr, bodies=bodies)


def half_fluid_cos(n=32):
    """The reference's MGPCG_Test problem (mgpcg.py:501-513,531-558): tag FLUID
    for j < n/2 and AIR above, r = cos(5 pi i/n) cos(5 pi j/n) cos(5 pi k/n) on
    FLUID cells and 0 on AIR."""
    i = np.arange(n, dtype=F32)
    c = np.cos(F32(5.0 * np.pi) * i / F32(n)).astype(F32)
    r = (c[:, None, None] * c[None, :, None] * c[None, None, :]).astype(F32)
    tag = np.zeros((n, n, n), np.int32)
    tag[:, : n // 2, :] = 1
    r[tag == 0] = 0.0
    return tag, r


SCENES = {
    "dam_break_64": lambda: dam_break(64, 0),
    "dam_break_spheres_128": lambda: dam_break_spheres(128, 1),
    "flow_past_sphere_256": lambda: flow_past_sphere(256, 2),
    "droplets_box_512": lambda: dam_break_droplets_box(512, 3),
}
