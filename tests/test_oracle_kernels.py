"""Pins of the CPU oracle's per-kernel restatements (no GPU).

Each check compares the oracle with something the mathematics fixes and that
is computed independently of the oracle's code: the exhaustive truth table of
restrict_tag (= min of the children), the sparse graph Laplacian of
tests/lapref.py (operator, divd, Gauss-Seidel), reshape sums (restriction),
np.repeat and adjointness (prolongation), and the divergence theorem /
gradient adjointness (divergence).  A dropped term, a wrong colour, sign or
index fails one of them.
"""
import itertools

import numpy as np
import pytest

from lapref import assemble, face_gradient, random_tags


# ------------------------------------------------------------------ tags
def test_restrict_tag_is_min_over_all_child_combinations(oracle):
    combos = np.array(list(itertools.product(range(3), repeat=8)), np.int32)   # 3^8 = 6561
    n = len(combos)
    fine = np.zeros((2, 2, 2 * n), np.int32)
    for idx, (ii, jj, kk) in enumerate(itertools.product(range(2), repeat=3)):
        fine[ii, jj, kk::2] = combos[:, idx]
    coarse = oracle.restrict_tag(fine)
    assert coarse.shape == (1, 1, n)
    assert np.array_equal(coarse[0, 0], combos.min(axis=1))


def test_tag_rules(oracle):
    phi = np.array([-1.0, -0.0, 0.0, 1e-30, 2.0, np.float32(-1e-30)], np.float32).reshape(1, 1, 6)
    t = oracle.tag_dirichlet(phi)
    assert t.ravel().tolist() == [1, 1, 1, 0, 0, 1]           # phi > 0 -> AIR, else FLUID
    rig = np.array([1.0, -1.0, 0.0, -0.5, 3.0, -2.0], np.float32).reshape(1, 1, 6)
    t2 = oracle.tag_neumann(rig, t)
    assert t2.ravel().tolist() == [1, 2, 1, 2, 0, 2]          # phi_rigid < 0 -> SOLID overrides


@pytest.mark.parametrize("shape,seed", [((6, 5, 4), 0), ((8, 8, 8), 1), ((3, 7, 10), 2)])
def test_divd_equals_matrix_diagonal(shape, seed, oracle):
    tag = random_tags(shape, seed)
    A, ids = assemble(tag)
    d = oracle.compute_divd(tag, np.float64)
    fl = tag == 1
    assert np.array_equal(d[fl], A.diagonal()[ids[fl]])
    assert np.all(d[~fl] == 0)


# ------------------------------------------------------------------ operator
@pytest.mark.parametrize("shape,seed", [((6, 5, 4), 3), ((8, 8, 8), 4), ((5, 9, 6), 5)])
def test_compute_Ap_is_minus_A(shape, seed, oracle):
    tag = random_tags(shape, seed)
    A, ids = assemble(tag)
    x = np.random.default_rng(seed).normal(size=shape)
    Ap = oracle.compute_Ap(x, tag, real=np.float64)
    fl = tag == 1
    ref = np.zeros(shape)
    ref[fl] = -(A @ x[fl])          # ids number FLUID cells in C order, as x[fl]
    assert np.allclose(Ap, ref, atol=1e-12)
    assert np.all(Ap[~fl] == 0)


def test_operator_symmetric_and_definite(oracle):
    tag = random_tags((7, 6, 5), 6)
    A, _ = assemble(tag)
    D = A.toarray()
    assert np.allclose(D, D.T)
    # each face-connected FLUID component touching AIR makes A SPD on it
    ev = np.linalg.eigvalsh(D)
    assert ev.min() > -1e-12


def test_closed_fluid_box_is_pure_neumann(oracle):
    tag = np.ones((6, 5, 4), np.int32)
    ones = np.ones(tag.shape)
    assert np.allclose(oracle.compute_Ap(ones, tag, real=np.float64), 0.0)
    # the tag-free operator of the "default" smoother is the same Laplacian
    x = np.random.default_rng(1).normal(size=(6, 6, 6))
    t6 = np.ones((6, 6, 6), np.int32)
    assert np.allclose(oracle.compute_Ap(x, None, real=np.float64), oracle.compute_Ap(x, t6, real=np.float64),
                       atol=1e-12)


# ------------------------------------------------------------------ smoother
@pytest.mark.parametrize("phase", [0, 1])
def test_smoother_is_textbook_gauss_seidel_on_one_colour(phase, oracle):
    shape = (8, 6, 6)
    tag = random_tags(shape, 7, p=(0.15, 0.75, 0.1))
    A, ids = assemble(tag)
    rng = np.random.default_rng(8)
    z = rng.normal(size=shape)
    z[tag != 1] = 0.0
    r = rng.normal(size=shape)
    divd = oracle.compute_divd(tag, np.float64)
    got = oracle.smooth(z, r, tag, divd, phase, real=np.float64)
    # textbook GS on the colour (i+j+k)&1 == phase^1: z_I = (r_I - sum_{J!=I} A_IJ z_J) / A_II
    fl = tag == 1
    zf = z[fl].copy()
    D = A.diagonal()
    i, j, k = np.nonzero(fl)
    colour = ((i + j + k) & 1) == (phase ^ 1)
    off = A @ zf - D * zf
    new = zf.copy()
    upd = colour & (D != 0)
    new[upd] = (r[fl][upd] - off[upd]) / D[upd]
    new[colour & (D == 0)] = 0.0
    exp = np.zeros(shape)
    exp[fl] = new
    assert np.allclose(got, exp, atol=1e-12)


# ------------------------------------------------------------------ transfer
def test_restriction_is_blocked_residual_sum(oracle):
    shape = (8, 8, 8)
    tag = random_tags(shape, 9, p=(0.1, 0.85, 0.05))
    tagc = oracle.restrict_tag(tag)
    A, ids = assemble(tag)
    rng = np.random.default_rng(10)
    z = rng.normal(size=shape)
    r = rng.normal(size=shape)
    r[tag != 1] = 0.0
    got = oracle.restrict(r, z, tag, tagc, real=np.float64)
    res = r.copy()
    fl = tag == 1
    res[fl] = r[fl] - (A @ z[fl])
    blocks = res.reshape(4, 2, 4, 2, 4, 2).sum(axis=(1, 3, 5))
    exp = np.where(tagc == 1, blocks, 0.0)
    assert np.allclose(got, exp, atol=1e-12)


def test_restriction_gather_equals_serial_atomic_order_bitwise(oracle):
    shape = (8, 12, 16)
    tag = random_tags(shape, 11, p=(0.1, 0.8, 0.1))
    tagc = oracle.restrict_tag(tag)
    rng = np.random.default_rng(12)
    z = rng.normal(size=shape).astype(np.float32)
    r = rng.normal(size=shape).astype(np.float32)
    a = oracle.restrict(r, z, tag, tagc, serial=True)
    b = oracle.restrict(r, z, tag, tagc, serial=False)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_prolongation_is_injection_and_adjoint_of_sum_restriction(oracle):
    rng = np.random.default_rng(13)
    zf = rng.normal(size=(8, 6, 4))
    zc = rng.normal(size=(4, 3, 2))
    got = oracle.prolongate(zf, zc, real=np.float64)
    assert np.allclose(got, zf + np.repeat(np.repeat(np.repeat(zc, 2, 0), 2, 1), 2, 2))
    # <R f, c> = <f, P c> with R the (tag-free, z = 0) sum restriction
    f = rng.normal(size=(8, 6, 4))
    Rf = oracle.restrict(f, np.zeros_like(f), real=np.float64)
    Pc = oracle.prolongate(np.zeros_like(f), zc, real=np.float64)
    assert np.isclose(np.sum(Rf * zc), np.sum(f * Pc))


# ------------------------------------------------------------------ divergence / gradient
def _random_mac(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.normal(size=(n[0] + 1, n[1], n[2])), rng.normal(size=(n[0], n[1] + 1, n[2])),
            rng.normal(size=(n[0], n[1], n[2] + 1)))


def test_divergence_is_minus_adjoint_of_gradient(oracle):
    n = (6, 6, 6)
    u, v, w = _random_mac(n, 14)
    z = [np.zeros_like(a) for a in (u, v, w)]
    tag = np.ones(n, np.int32)
    F = oracle.divergence_tag(u, v, w, *z, tag, real=np.float64)
    p = np.random.default_rng(15).normal(size=n)
    gu, gv, gw = face_gradient(p)
    assert np.isclose(np.sum(F * p), -(np.sum(u * gu) + np.sum(v * gv) + np.sum(w * gw)))
    # divergence theorem with closed walls: the fluxes telescope to zero
    assert abs(F.sum()) < 1e-10


def test_divergence_solid_override_equals_substituted_field(oracle):
    n = (6, 6, 6)
    u, v, w = _random_mac(n, 16)
    ur, vr, wr = _random_mac(n, 17)
    tag = random_tags(n, 18, p=(0.2, 0.6, 0.2))
    F = oracle.divergence_tag(u, v, w, ur, vr, wr, tag, real=np.float64)
    so = tag == 2
    us, vs, ws = u.copy(), v.copy(), w.copy()
    # a face whose cell on the far side is SOLID takes the rigid velocity
    us[1:-1][so[:-1] | so[1:]] = ur[1:-1][so[:-1] | so[1:]]
    vs[:, 1:-1][so[:, :-1] | so[:, 1:]] = vr[:, 1:-1][so[:, :-1] | so[:, 1:]]
    ws[:, :, 1:-1][so[:, :, :-1] | so[:, :, 1:]] = wr[:, :, 1:-1][so[:, :, :-1] | so[:, :, 1:]]
    ref = (us[1:] - us[:-1]) + (vs[:, 1:] - vs[:, :-1]) + (ws[:, :, 1:] - ws[:, :, :-1])
    # wall faces are zero in the reference kernel
    ref[0] += us[0]; ref[-1] -= us[-1]
    ref[:, 0] += vs[:, 0]; ref[:, -1] -= vs[:, -1]
    ref[:, :, 0] += ws[:, :, 0]; ref[:, :, -1] -= ws[:, :, -1]
    # the override keys on the neighbour across the face; for a non-SOLID cell
    # that is exactly "either side of the face is SOLID"
    m = tag != 2
    assert np.allclose(F[m], ref[m], atol=1e-12)


@pytest.mark.parametrize("d", [1, 2, 3])
def test_subtract_gradient_face_rules(d, oracle):
    n = (5, 6, 7)
    tag = random_tags(n, 19)
    rng = np.random.default_rng(20)
    p = rng.normal(size=n)
    shape = tuple(s + (a == d - 1) for a, s in enumerate(n))
    uf = rng.normal(size=shape)
    ur = rng.normal(size=shape)
    out = oracle.subtract_gradient_tag(uf, ur, p, tag, d, real=np.float64)
    ax = d - 1
    lo = [slice(None)] * 3
    hi = [slice(None)] * 3
    lo[ax] = slice(0, -1)
    hi[ax] = slice(1, None)
    tl, tr = tag[tuple(lo)], tag[tuple(hi)]
    inner = [slice(None)] * 3
    inner[ax] = slice(1, -1)
    got = out[tuple(inner)]
    exp = np.where((tl == 2) | (tr == 2), ur[tuple(inner)],
                   np.where((tl == 1) | (tr == 1), uf[tuple(inner)] - (p[tuple(hi)] - p[tuple(lo)]), 0.0))
    assert np.allclose(got, exp)
    first = [slice(None)] * 3
    last = [slice(None)] * 3
    first[ax] = 0
    last[ax] = -1
    assert np.all(out[tuple(first)] == 0) and np.all(out[tuple(last)] == 0)


def test_zero_centered(oracle):
    tag = random_tags((5, 5, 5), 21)
    q = np.random.default_rng(22).normal(size=(5, 5, 5))
    out = oracle.zero_centered(q, tag, real=np.float64)
    fl = tag == 1
    assert abs(out[fl].mean()) < 1e-12
    assert np.all(out[~fl] == 0)
    assert np.allclose(out[fl] - q[fl], -q[fl].mean())
