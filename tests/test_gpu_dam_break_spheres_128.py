"""Plain FP32 MGPCG at 128^3 on the dam-break block union 32 fluid spheres
(radius 2-8 cells, centres uniform, seed 1), fluid-face velocities N(0, 0.5)
and zero on every other face (``scenes.dam_break_spheres(128, 1)``), against
the CPU oracle through the C-ABI:

* cell flags (geometry.py:335-344, :390-397) and the coarse masks and divd of
  every level (mgpcg.py:145-189 set_tags) bit-exact;
* the divergence (divergence.py:44-95, masked and scaled) bitwise;
* the V-cycle the fused solve runs (colour-split level 0) bitwise;
* the projection (example_fluid.py:120-147): iterations within one of the FP32
  oracle's, pressure and velocities within 1e-5 of it, pressure at least as
  close to the converged FP64 solution as the FP32 reference algorithm, and
  the post-projection divergence small.
"""
import ctypes

import numpy as np
import pytest

from flub_amd import _lib, scenes
from flub_amd import device as wp
from flub_amd.const_param import Field, IntField
from flub_amd.geometry import tag_dirichlet, tag_neumann
from flub_amd.multigrid.mgpcg import MGPCG
from flub_amd.project import Projection

pytestmark = pytest.mark.gpu

N, LEVELS = 128, 4


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _rel(a, b):
    return float(np.linalg.norm((np.asarray(a, np.float64) - b).ravel()) / np.linalg.norm(np.ravel(b)))


@pytest.fixture(scope="module")
def scene():
    s = scenes.dam_break_spheres(N, 1)
    assert s["n"] == N
    return s


def test_scene_is_the_described_one(scene):
    """32 spheres joined to the dam-break block: more FLUID than the block
    alone, AIR elsewhere, no SOLID cell (walls are the clamped boundary)."""
    tag = scene["tag"]
    block = scenes.dam_break(N, 0)["tag"]
    assert (tag == 2).sum() == 0
    assert (tag == 1).sum() > (block == 1).sum()
    assert (tag[block == 1] == 1).all()
    fl = tag == 1
    assert np.abs(scene["u"]).max() > 0 and 0.4 < float(np.std(scene["u"][1:-1][fl[:-1] & fl[1:]])) < 0.6


def test_flags_and_coarse_masks_bitexact(scene, oracle):
    phi = Field(f=wp.from_numpy(scene["phi"]))
    phr = Field(f=wp.from_numpy(scene["phi_rigid"]))
    tag = IntField(f=wp.zeros(scene["phi"].shape, int))
    tag_dirichlet(tag, phi)
    tag_neumann(tag, phr)
    t = oracle.tag_neumann(scene["phi_rigid"], oracle.tag_dirichlet(scene["phi"]))
    assert np.array_equal(tag.f.numpy(), t)
    assert np.array_equal(t, scene["tag"])
    for levels in (LEVELS, 6):
        mg = MGPCG(3, (N,) * 3, levels)
        mg.set_tags(tag.f)
        tl = t
        for l in range(levels):
            assert np.array_equal(mg.tag[l].numpy().astype(np.int32), tl), f"coarse mask level {l}"
            assert bits_equal(mg.divd[l].numpy(), oracle.compute_divd(tl)), f"divd level {l}"
            if l + 1 < levels:
                tl = oracle.restrict_tag(tl)


def test_split_vcycle_bitexact(scene, oracle):
    tag = scene["tag"]
    r = np.random.default_rng(31).normal(0, 1, tag.shape).astype(np.float32)
    r[tag != 1] = 0.0
    mg = MGPCG(3, tag.shape, LEVELS)
    mg.pre_and_post_smoothing, mg.bottom_smoothing = 2, 6
    mg.set_tags(wp.from_numpy(tag))
    rd, zd = wp.from_numpy(r), wp.zeros(tag.shape)
    _lib.call("flub_mgpcg_precondition_into", mg.handle, ctypes.c_void_p(rd.ptr), ctypes.c_void_p(zd.ptr), 1)
    orc = oracle.MGPCG(tag.shape, LEVELS, pre_and_post_smoothing=2, bottom_smoothing=6)
    orc.set_tags(tag)
    orc.put("r", r, 0)
    orc.apply_preconditioner("phases")
    assert bits_equal(zd.numpy(), orc.get("z", 0))


def test_projection_matches_oracle(scene, oracle):
    s = scene
    pr = Projection((N,) * 3, n_mg_levels=LEVELS, pre_and_post_smoothing=2, bottom_smoothing=6)
    pr.load(s)
    it = pr.project(max_iters=-1, rel_tol=1e-12)
    p_gpu = pr.p0.f.numpy()
    u, v, w = pr.velocity()
    o32 = oracle.MGPCG((N,) * 3, LEVELS, pre_and_post_smoothing=2, bottom_smoothing=6)
    ru, rv, rw, rp, rdiv, rit, _ = o32.project(s["u"], s["v"], s["w"], s["ur"], s["vr"], s["wr"], s["tag"],
                                               1.0 / N, rel_tol=1e-12)
    del o32
    o64 = oracle.MGPCG((N,) * 3, LEVELS, real=np.float64, pre_and_post_smoothing=2, bottom_smoothing=6)
    _, _, _, p64, _, _, _ = o64.project(s["u"], s["v"], s["w"], s["ur"], s["vr"], s["wr"], s["tag"], 1.0 / N,
                                        rel_tol=1e-24)
    del o64
    e_gpu, e_ref = _rel(p_gpu, p64), _rel(rp, p64)
    print(f"128^3 dam break + 32 spheres: {it} iterations (oracle {rit}); pressure vs FP32 oracle "
          f"{_rel(p_gpu, rp):.2e}, vs FP64 {e_gpu:.2e} (FP32 oracle {e_ref:.2e})")
    assert bits_equal(pr.div.f.numpy(), rdiv)
    assert abs(it - rit) <= 1
    assert _rel(p_gpu, rp) < 1e-5
    assert e_gpu < 1.5 * e_ref + 1e-6
    assert _rel(u, ru) < 1e-5 and _rel(v, rv) < 1e-5 and _rel(w, rw) < 1e-5
    fl = s["tag"] == 1
    post = oracle.divergence_tag(u, v, w, s["ur"], s["vr"], s["wr"], s["tag"], real=np.float64)
    pre = oracle.divergence_tag(s["u"], s["v"], s["w"], s["ur"], s["vr"], s["wr"], s["tag"], real=np.float64)
    assert np.linalg.norm(post[fl]) < 2e-5 * np.linalg.norm(pre[fl])
