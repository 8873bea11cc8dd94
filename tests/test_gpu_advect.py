"""GPU parity of the advection kernels (advect.py:11-137) against the CPU
oracle, bitwise: the fused MacCormack kernel against the reference's three
launches restated literally (u, v, w and a centred field, small and large
time steps so that traces leave the grid and the clamps act), advect_phi and
advect_simple_3d."""
import numpy as np
import pytest

from flub_amd import device as wp
from flub_amd.advect import advect_3d, advect_mc, advect_phi, advect_simple_3d
from flub_amd.const_param import Field, MACGrid
from flub_amd.launch import launch

pytestmark = pytest.mark.gpu

STAG = {0: (0.5, 0.5, 0.5), 1: (0.0, 0.5, 0.5), 2: (0.5, 0.0, 0.5), 3: (0.5, 0.5, 0.0)}


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _vel(shape, seed, sigma):
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    return tuple(rng.normal(0, sigma, s).astype(np.float32)
                 for s in ((nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)))


def _mac(arrs):
    return MACGrid(*(Field(f=wp.from_numpy(a), stagger=STAG[d], d=d) for d, a in zip((1, 2, 3), arrs)))


@pytest.mark.parametrize("shape,seed,dt", [((32, 32, 32), 0, 0.01), ((48, 40, 32), 1, 0.01), ((32, 32, 32), 2, 0.2)])
def test_advect_mc_bitexact(shape, seed, dt, oracle):
    vel = _vel(shape, seed, 0.5)
    h0 = 1.0 / shape[0]
    V = _mac(vel)
    out = MACGrid(*(Field(f=wp.empty(a.shape, float), stagger=STAG[d], d=d) for d, a in zip((1, 2, 3), vel)))
    rng = np.random.default_rng(seed + 10)
    q = rng.normal(size=shape).astype(np.float32)
    Q = Field(f=wp.from_numpy(q), stagger=STAG[0], d=0)
    Qn = Field(f=wp.empty(shape, float), stagger=STAG[0], d=0)
    advect_3d(V, Q, out, Qn, dt=dt, h0=h0)
    assert np.array_equal(bits(Qn.f.numpy()), bits(oracle.advect_mc(*vel, q, STAG[0], dt, h0)))
    for d, a in zip((1, 2, 3), vel):
        want = oracle.advect_mc(*vel, a, STAG[d], dt, h0)
        assert np.array_equal(bits((out.u, out.v, out.w)[d - 1].f.numpy()), bits(want)), d


@pytest.mark.parametrize("dt", [0.01, 0.3])
def test_advect_phi_and_simple_bitexact(dt, oracle):
    shape = (32, 24, 40)
    vel = _vel(shape, 5, 0.7)
    h0 = 1.0 / shape[0]
    V = _mac(vel)
    rng = np.random.default_rng(6)
    phi = rng.normal(size=shape).astype(np.float32)
    P = Field(f=wp.from_numpy(phi), stagger=STAG[0], d=0)
    Pn = Field(f=wp.empty(shape, float), stagger=STAG[0], d=0)
    launch(advect_phi, dim=shape, inputs=[V, P, Pn, dt, h0])
    got = Pn.f.numpy()
    assert np.array_equal(bits(got), bits(oracle.advect_phi(*vel, phi, dt, h0)))
    if dt > 0.1:
        assert (got == np.float32(h0)).any()                 # traces that left the grid
    qn = wp.empty(shape, float)
    launch(advect_simple_3d, dim=shape, inputs=[V.u.f, V.v.f, V.w.f, P.f, qn, dt, h0])
    assert np.array_equal(bits(qn.numpy()), bits(oracle.advect_simple(*vel, phi, dt, h0)))


def test_advect_mc_rejects_in_place():
    vel = _vel((8, 8, 8), 7, 0.1)
    V = _mac(vel)
    with pytest.raises(Exception):
        advect_mc(V, V.u, V.u)
