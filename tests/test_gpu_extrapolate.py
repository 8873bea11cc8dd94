"""GPU parity of the velocity extrapolation (extrapolate.py:9-47,
utils/util.py:424-437, water_3d.py:70-73) against the CPU oracle, bitwise:
set_unknown (the last face layer keeps the owner values it held), one
extrapolate_3d launch + flush per axis, extrapolate_by_1, and the driver's
n passes (computed as one) against the oracle running every pass."""
import numpy as np
import pytest

from flub_amd import device as wp
from flub_amd import scenes
from flub_amd.const_param import Field, IntField, MACGrid, MACGridInt
from flub_amd.extrapolate import extrapolate, extrapolate_3d, extrapolate_by_1, flush_1_3d
from flub_amd.launch import launch
from flub_amd.utils.util import set_unknown

pytestmark = pytest.mark.gpu


def _mac(arrs, cls, fcls):
    return cls(*(fcls(f=wp.from_numpy(a), d=d) for d, a in zip((1, 2, 3), arrs)))


def _inputs(n, seed):
    s = scenes.dam_break_spheres(n, seed, n_spheres=6, r_cells=(2, 6))
    rng = np.random.default_rng(seed)
    vel = (s["u"], s["v"], s["w"])
    # owner arrays as the face fractions leave them: -1, or a body index
    owners = tuple(rng.choice([-1, 0, 1], size=a.shape, p=[0.8, 0.1, 0.1]).astype(np.int32) for a in vel)
    return s["phi"], vel, owners


@pytest.mark.parametrize("n,seed", [(32, 0), (64, 1)])
def test_set_unknown_and_extrapolate_by_1_bitexact(n, seed, oracle):
    phi, vel, owners = _inputs(n, seed)
    tags = oracle.set_unknown(phi, *owners)
    ref_v, ref_t = oracle.extrapolate_by_1(vel, tags)
    V = _mac(vel, MACGrid, Field)
    T = _mac(owners, MACGridInt, IntField)
    launch(set_unknown, dim=phi.shape, inputs=[Field(f=wp.from_numpy(phi)), T])
    for got, want in zip((T.u, T.v, T.w), tags):
        assert np.array_equal(got.f.numpy(), want)
    extrapolate_by_1(V, T)
    for got, want in zip((V.u, V.v, V.w), ref_v):
        assert np.array_equal(got.f.numpy().view(np.uint32), want.view(np.uint32))
    for got, want in zip((T.u, T.v, T.w), ref_t):
        assert np.array_equal(got.f.numpy(), want)


def test_single_launches_bitexact(oracle):
    phi, vel, owners = _inputs(48, 2)
    tags = oracle.set_unknown(phi, *owners)
    V = _mac(vel, MACGrid, Field)
    T = _mac(tags, MACGridInt, IntField)
    for d, (q, t) in enumerate(zip(vel, tags), start=1):
        rq, rt = oracle.extrapolate_3d(q, t)
        uf = (V.u, V.v, V.w)[d - 1]
        tf = (T.u, T.v, T.w)[d - 1]
        launch(extrapolate_3d, dim=uf.f.shape, inputs=[V, T, d])
        assert np.array_equal(uf.f.numpy().view(np.uint32), rq.view(np.uint32))
        assert np.array_equal(tf.f.numpy(), rt)
        assert (rt == -1).any()
        launch(flush_1_3d, dim=tf.f.shape, inputs=[tf.f])
        assert np.array_equal(tf.f.numpy(), oracle.flush_1(rt))


def test_driver_equals_every_pass(oracle):
    n = 64
    phi, vel, owners = _inputs(n, 3)
    tags = oracle.set_unknown(phi, *owners)
    ref_v, ref_t = vel, tags
    for _ in range(16 * n // 128):
        ref_v, ref_t = oracle.extrapolate_by_1(ref_v, ref_t)
    V = _mac(vel, MACGrid, Field)
    T = _mac(owners, MACGridInt, IntField)
    extrapolate(V, T, Field(f=wp.from_numpy(phi)))
    for got, want in zip((V.u, V.v, V.w), ref_v):
        assert np.array_equal(got.f.numpy().view(np.uint32), want.view(np.uint32))
    for got, want in zip((T.u, T.v, T.w), ref_t):
        assert np.array_equal(got.f.numpy(), want)


@pytest.mark.parametrize("passes", [0, 1, 2, 3])
def test_driver_pass_counts(passes, oracle):
    """Zero and one pass, and two or more computed as one launch per axis,
    against the oracle running every pass."""
    n = 32
    phi, vel, owners = _inputs(n, 5)
    ref_v, ref_t = vel, oracle.set_unknown(phi, *owners)
    for _ in range(passes):
        ref_v, ref_t = oracle.extrapolate_by_1(ref_v, ref_t)
    V = _mac(vel, MACGrid, Field)
    T = _mac(owners, MACGridInt, IntField)
    extrapolate(V, T, Field(f=wp.from_numpy(phi)), n_cells=passes)
    for got, want in zip((V.u, V.v, V.w), ref_v):
        assert np.array_equal(got.f.numpy().view(np.uint32), np.asarray(want).view(np.uint32))
    for got, want in zip((T.u, T.v, T.w), ref_t):
        assert np.array_equal(got.f.numpy(), want)


def _inputs_shape(shape, seed):
    """Non-cubic grids: a level set of random spheres (in cell units) on an
    nx x ny x nz grid, random face velocities and owner marks."""
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    i, j, k = np.meshgrid(np.arange(nx) + 0.5, np.arange(ny) + 0.5, np.arange(nz) + 0.5, indexing="ij")
    phi = np.full(shape, np.inf, np.float32)
    for _ in range(6):
        c = rng.uniform(0, 1, 3) * np.array(shape)
        r = rng.uniform(3, 9)
        phi = np.minimum(phi, (np.sqrt((i - c[0]) ** 2 + (j - c[1]) ** 2 + (k - c[2]) ** 2) - r).astype(np.float32))
    vel = (rng.normal(0, 1, (nx + 1, ny, nz)).astype(np.float32), rng.normal(0, 1, (nx, ny + 1, nz)).astype(np.float32),
           rng.normal(0, 1, (nx, ny, nz + 1)).astype(np.float32))
    owners = tuple(rng.choice([-1, 0, 1], size=a.shape, p=[0.8, 0.1, 0.1]).astype(np.int32) for a in vel)
    return phi, vel, owners


@pytest.mark.parametrize("shape,seed", [((40, 24, 56), 6), ((72, 40, 56), 7), ((17, 9, 70), 8)])
def test_extrapolate_by_1_non_cubic_bitexact(shape, seed, oracle):
    """The LDS-tiled fill (8 x 8 x 62 face tiles) on grids whose extents are not
    multiples of the tile, and not equal to each other."""
    phi, vel, owners = _inputs_shape(shape, seed)
    tags = oracle.set_unknown(phi, *owners)
    ref_v, ref_t = oracle.extrapolate_by_1(vel, tags)
    V = _mac(vel, MACGrid, Field)
    T = _mac(owners, MACGridInt, IntField)
    launch(set_unknown, dim=phi.shape, inputs=[Field(f=wp.from_numpy(phi)), T])
    extrapolate_by_1(V, T)
    for got, want in zip((V.u, V.v, V.w), ref_v):
        assert np.array_equal(got.f.numpy().view(np.uint32), want.view(np.uint32))
    for got, want in zip((T.u, T.v, T.w), ref_t):
        assert np.array_equal(got.f.numpy(), want)


@pytest.mark.parametrize("shape,passes", [((40, 24, 56), 3), ((72, 40, 56), 8)])
def test_driver_non_cubic_equals_every_pass(shape, passes, oracle):
    phi, vel, owners = _inputs_shape(shape, 9)
    ref_v, ref_t = vel, oracle.set_unknown(phi, *owners)
    for _ in range(passes):
        ref_v, ref_t = oracle.extrapolate_by_1(ref_v, ref_t)
    V = _mac(vel, MACGrid, Field)
    T = _mac(owners, MACGridInt, IntField)
    extrapolate(V, T, Field(f=wp.from_numpy(phi)), n_cells=passes)
    for got, want in zip((V.u, V.v, V.w), ref_v):
        assert np.array_equal(got.f.numpy().view(np.uint32), np.asarray(want).view(np.uint32))
    for got, want in zip((T.u, T.v, T.w), ref_t):
        assert np.array_equal(got.f.numpy(), want)
