"""The split solves' p step and Ap in one pass (csrc/split.hip, k_sp_pAp:
p = beta p + z, then Ap = ndt(p) and p.Ap) against the FP32 oracle's
compute_Ap and against the two-pass step (FLUB_NO_PAP=1: k_update_pp, then
k_sp_Ap), through flub_mgpcg_profile's p-step entry (kernel 14: p of the
handle and the split z of the last V-cycle, Ap into the handle's Ap, p.Ap
into the `sum` scalar):

* Ap is bitwise the oracle's Ap of beta p + z (beta 0.5, so that beta p + z
  rounds once in numpy as in the kernel's fused multiply-add);
* for any beta, Ap is bitwise that of the two-pass step and p.Ap agrees to
  the FP64 rounding of another summation grouping;
* grids whose segments along i end short of a full segment, with blocks past
  the last tile (52 rows by 8), and one to four tiles of j rows;
* whole FP32 and BF16 solves with the one-pass step take the iterations of the
  two-pass step and agree to 1e-6 in pressure.
"""
import ctypes

import numpy as np
import pytest

from flub_amd import _lib, scenes
from flub_amd import device as wp
from flub_amd.multigrid.mgpcg import MGPCG
from flub_amd.project import Projection

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def split_index(shape):
    """Position of natural cell (i, j, k) in the colour-split layout: colour
    (i + j + k) & 1 selects the half, ((i * ny + j) * nz / 2 + k / 2) the place."""
    nx, ny, nz = shape
    i, j, k = np.indices(shape)
    return (((i + j + k) & 1) * (nx * ny * nz // 2) + (i * ny + j) * (nz // 2) + k // 2).ravel()


def to_split(a):
    out = np.empty(a.size, a.dtype)
    out[split_index(a.shape)] = a.ravel()
    return out


def from_split(s, shape):
    return np.asarray(s).ravel()[split_index(shape)].reshape(shape)


def _tags(shape, seed):
    if len(set(shape)) == 1:
        return scenes.dam_break_spheres(shape[0], seed, n_spheres=8, r_cells=(2, 5))["tag"]
    rng = np.random.default_rng(seed)
    return rng.choice(3, size=shape, p=[0.1, 0.8, 0.1]).astype(np.int32)


def _pstep(tag, r, p, beta):
    """The handle's p step: returns (Ap in the natural layout, p.Ap, the V-cycle's z)."""
    mg = MGPCG(3, tag.shape, 3)
    mg.pre_and_post_smoothing, mg.bottom_smoothing = 2, 6
    mg.set_tags(wp.from_numpy(tag))
    rd, zd = wp.from_numpy(r), wp.zeros(tag.shape)
    _lib.call("flub_mgpcg_precondition_into", mg.handle, ctypes.c_void_p(rd.ptr), ctypes.c_void_p(zd.ptr), 1)
    mg.p.assign(to_split(p))
    mg.beta.assign(np.array([beta], np.float64))
    avg, byt = ctypes.c_float(), ctypes.c_double()
    _lib.call("flub_mgpcg_profile", mg.handle, 14, 1, ctypes.byref(avg), ctypes.byref(byt))
    wp.synchronize()
    return from_split(mg.Ap.numpy(), tag.shape), float(mg.sum.numpy()[0]), zd.numpy()


@pytest.mark.parametrize("shape,ib", [((64, 64, 64), None), ((128, 128, 128), None), ((52, 64, 128), "8"),
                                      ((96, 32, 256), "4"), ((256, 256, 256), None)])
def test_fused_pstep_ap(shape, ib, oracle, monkeypatch):
    tag = _tags(shape, 5)
    rng = np.random.default_rng(29)
    r = rng.normal(0, 1, shape).astype(np.float32)
    r[tag != 1] = 0.0
    p = rng.normal(0, 1, shape).astype(np.float32)
    if ib:
        monkeypatch.setenv("FLUB_PAP_IB", ib)
    ap_half, s_half, z = _pstep(tag, r, p, 0.5)
    p_new = (np.float32(0.5) * p + z).astype(np.float32)
    want = oracle.compute_Ap(p_new, tag)
    assert bits_equal(ap_half, want)
    ref_dot = float(np.dot(p_new.astype(np.float64).ravel(), want.astype(np.float64).ravel()))
    assert abs(s_half - ref_dot) <= 1e-12 * (abs(ref_dot) + np.abs(p_new).sum())
    # any beta: bitwise the two-pass step
    ap_fused, s_fused, _ = _pstep(tag, r, p, 0.3)
    monkeypatch.setenv("FLUB_NO_PAP", "1")
    ap_two, s_two, _ = _pstep(tag, r, p, 0.3)
    assert bits_equal(ap_fused, ap_two)
    assert abs(s_fused - s_two) <= 1e-12 * abs(s_two)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_fused_pstep_solve(precision, monkeypatch):
    s = scenes.flow_past_sphere(128, 2)
    n = s["n"]
    res = {}
    for mode in ("fused", "two-pass"):
        if mode == "two-pass":
            monkeypatch.setenv("FLUB_NO_PAP", "1")
        pr = Projection((n,) * 3, n_mg_levels=4, pre_and_post_smoothing=2, bottom_smoothing=6, precision=precision)
        pr.load(s)
        it = pr.project(max_iters=-1, rel_tol=1e-12)
        res[mode] = (it, pr.p0.f.numpy().astype(np.float64))
        del pr
    print(f"{precision}: one pass {res['fused'][0]} iterations, two passes {res['two-pass'][0]}")
    assert res["fused"][0] == res["two-pass"][0]
    d = np.linalg.norm(res["fused"][1] - res["two-pass"][1]) / np.linalg.norm(res["two-pass"][1])
    assert d < 1e-6
