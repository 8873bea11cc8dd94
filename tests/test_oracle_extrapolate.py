"""Pins of the oracle's velocity extrapolation (extrapolate.py:9-47,
utils/util.py:424-437) by independent computations:

* a vectorised NumPy mean over the known (tag 0) faces of each unknown face's
  in-range 27-neighbourhood, in FP64;
* constant fields stay exactly constant;
* the reference's flush turns the -1 marks back into 1, so with marks in
  {0, 1} a full extrapolate_by_1 leaves them as they were and repeating it
  changes nothing; with -1 marks present (owner values left in the last face
  layer) the second pass differs from the first and the third equals the
  second (the driver's 16 * res / 128 passes equal two);
* set_unknown writes only the first nx x ny x nz entries of each face array.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _mean_known(q, unknown):
    """FP64 mean over the in-range 27-neighbourhood faces with tag 0."""
    sx, sy, sz = q.shape
    pad_q = np.zeros((sx + 2, sy + 2, sz + 2))
    pad_k = np.zeros((sx + 2, sy + 2, sz + 2))
    pad_q[1:-1, 1:-1, 1:-1] = q
    pad_k[1:-1, 1:-1, 1:-1] = unknown == 0
    s = np.zeros(q.shape)
    n = np.zeros(q.shape)
    for a in range(3):
        for b in range(3):
            for c in range(3):
                k = pad_k[a:a + sx, b:b + sy, c:c + sz]
                s += k * pad_q[a:a + sx, b:b + sy, c:c + sz]
                n += k
    return s, n


@pytest.mark.parametrize("shape,seed", [((9, 8, 7), 0), ((17, 16, 16), 1)])
def test_extrapolate_is_mean_of_known_neighbours(shape, seed):
    rng = np.random.default_rng(seed)
    q = rng.normal(size=shape).astype(np.float32)
    unknown = rng.choice([0, 1], size=shape, p=[0.6, 0.4]).astype(np.int32)
    out, t = O.extrapolate_3d(q, unknown)
    s, n = _mean_known(q.astype(np.float64), unknown)
    unk = unknown == 1
    has = unk & (n > 0)
    np.testing.assert_allclose(out[has], (s / np.maximum(n, 1))[has], rtol=1e-6, atol=1e-6)
    assert (out[unk & (n == 0)] == 0.0).all()
    assert np.array_equal(out[~unk], q[~unk])               # known faces untouched
    assert (t[has] == -1).all() and (t[unk & (n == 0)] == 1).all() and np.array_equal(t[~unk], unknown[~unk])


def test_constant_field_and_idempotence():
    rng = np.random.default_rng(2)
    shapes = [(9, 8, 8), (8, 9, 8), (8, 8, 9)]
    vel = [np.full(s, 1.25, np.float32) for s in shapes]
    tags = [rng.choice([0, 1], size=s, p=[0.5, 0.5]).astype(np.int32) for s in shapes]
    v1, t1 = O.extrapolate_by_1(vel, tags)
    for a, t0, t in zip(v1, tags, t1):
        _, n = _mean_known(np.zeros(a.shape), t0)
        assert (a[(t0 == 0) | (n > 0)] == 1.25).all()
        assert np.array_equal(t, t0)                        # the flush restores every mark
    v2, t2 = O.extrapolate_by_1(v1, t1)
    for a, b in zip(v1, v2):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    # a random field: one pass fixes it, the next changes nothing
    vel = [rng.normal(size=s).astype(np.float32) for s in shapes]
    v1, t1 = O.extrapolate_by_1(vel, tags)
    v2, _ = O.extrapolate_by_1(v1, t1)
    for a, b in zip(v1, v2):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_set_unknown_layout():
    rng = np.random.default_rng(3)
    nx, ny, nz = 6, 5, 4
    phi = rng.normal(size=(nx, ny, nz)).astype(np.float32)
    tu = np.full((nx + 1, ny, nz), 7, np.int32)
    tv = np.full((nx, ny + 1, nz), 7, np.int32)
    tw = np.full((nx, ny, nz + 1), 7, np.int32)
    u, v, w = O.set_unknown(phi, tu, tv, tw)
    want = (phi >= 0).astype(np.int32)
    assert np.array_equal(u[:nx], want) and (u[nx] == 7).all()
    assert np.array_equal(v[:, :ny], want) and (v[:, ny] == 7).all()
    assert np.array_equal(w[:, :, :nz], want) and (w[:, :, nz] == 7).all()


def test_preexisting_minus_one_marks_need_two_passes():
    rng = np.random.default_rng(4)
    shapes = [(9, 8, 8), (8, 9, 8), (8, 8, 9)]
    vel = [rng.normal(size=s).astype(np.float32) for s in shapes]
    tags = [rng.choice([-1, 0, 1], size=s, p=[0.2, 0.5, 0.3]).astype(np.int32) for s in shapes]
    v1, t1 = O.extrapolate_by_1(vel, tags)
    v2, t2 = O.extrapolate_by_1(v1, t1)
    v3, t3 = O.extrapolate_by_1(v2, t2)
    for t0, a1, b1, a2, b2, a3, b3 in zip(tags, v1, t1, v2, t2, v3, t3):
        assert ((b1 == 0) | (b1 == 1)).all() and np.array_equal(b1, np.where(t0 == -1, 1, t0))
        assert not np.array_equal(a1, a2)                       # the former -1 faces filled
        assert np.array_equal(a2.view(np.uint32), a3.view(np.uint32)) and np.array_equal(b2, b3)
