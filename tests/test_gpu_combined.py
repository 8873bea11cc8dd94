"""The projection of a scene with a solid body at a random pose, end to end
on the GPU: face fractions and rigid face velocities from the body, the
fraction-weighted divergence (divergence.py:5-41) as right-hand side, the
pressure solve, the gradient subtraction.

The inputs of the solve are checked bitwise against the CPU oracle's chain
(face fractions -> solid velocities -> fractional divergence -> mask ->
scale).  The pressure is checked against the converged FP64 oracle solve of
that right-hand side, for the plain FP32 MGPCG and for the combined learned
(U-Net) + BF16 mixed-precision solver, to the accuracy of the FP32 reference
algorithm itself.
"""
import os

import numpy as np
import pytest

from flub_amd import scenes
from flub_amd.project import Projection

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def oracle_chain(oracle, s):
    n = s["n"]
    h = np.float32(1.0 / n)
    b = s["bodies"]
    rb = oracle.Rigids(b["x"], b["q"], b["gtype"], b["radius"], b["v"], b["omega"])
    (fu, fv, fw), (ou, ov, ow) = oracle.face_fractions((n, n, n), h, rb)
    ur, vr, wr = (oracle.solid_vel(o, d, h, rb) for d, o in ((1, ou), (2, ov), (3, ow)))
    div = oracle.divergence_fractional(s["u"], s["v"], s["w"], fu, fv, fw, ur, vr, wr)
    div = oracle.mask_with_tag(div, s["tag"])
    div = oracle.scale(div, -1.0 / (1.0 / n))
    return div, (ur, vr, wr)


def oracle_solve(oracle, tag, div, n, real, rel_tol, eps):
    o = oracle.MGPCG((n, n, n), 4, real=real, pre_and_post_smoothing=2, bottom_smoothing=6)
    o.set_tags(tag)
    o.init(div.astype(real), -1.0 / n)
    o.mask_interior()
    it, _ = o.solve(max_iters=-1, eps=eps, abs_tol=0.0, rel_tol=rel_tol, smoother="phases")
    return o.get("x") * real(1.0 / n), it


@pytest.fixture(scope="module")
def scene():
    return scenes.dam_break_droplets_box(64, 3, n_drops=20)


@pytest.fixture(scope="module")
def reference(scene, oracle):
    n = scene["n"]
    div, rig = oracle_chain(oracle, scene)
    p64, _ = oracle_solve(oracle, scene["tag"], div.astype(np.float64), n, np.float64, 1e-24, 0.0)
    p32, it32 = oracle_solve(oracle, scene["tag"], div, n, np.float32, 1e-12, 1e-12)
    return div, rig, p64, p32, it32


def test_solid_scene_inputs_and_fp32_solve(scene, reference):
    div, rig, p64, p32, it32 = reference
    n = scene["n"]
    assert (scene["tag"] == 2).sum() > 0
    pr = Projection((n, n, n), n_mg_levels=4, pre_and_post_smoothing=2, bottom_smoothing=6)
    pr.load(scene)
    pr.set_bodies(scene["bodies"])
    it = pr.project_fractional(max_iters=-1, rel_tol=1e-12)
    for got, want in zip((pr.vel_rigid.u, pr.vel_rigid.v, pr.vel_rigid.w), rig):
        assert bits_equal(got.f.numpy(), want)
    assert bits_equal(pr.div.f.numpy(), div)
    assert abs(it - it32) <= 1
    p = pr.p0.f.numpy()
    err, err_ref = _rel(p, p64), _rel(p32, p64)
    assert _rel(p, p32) < 1e-5
    assert err < 1.5 * err_ref + 1e-6, (err, err_ref)


def test_combined_learned_mixed_solve(scene, reference):
    import torch
    from flub_amd.learned.evaluate import load_net
    from flub_amd.learned.fpcg import LearnedPCG

    div, rig, p64, p32, it32 = reference
    n = scene["n"]
    net = load_net(os.path.join(ROOT, "flub_amd", "learned", "weights", "unet_128.npz"))
    pr = Projection((n, n, n), n_mg_levels=4, pre_and_post_smoothing=2, bottom_smoothing=6, precision="bf16")
    pr.load(scene)
    pr.set_bodies(scene["bodies"])
    fluid = torch.from_numpy(scene["tag"] == 1).cuda()
    st = pr.project_fractional(LearnedPCG(pr.mgpcg, net), fluid, rel_tol=1e-12)
    assert bits_equal(pr.div.f.numpy(), div)
    assert st["net_iters"] >= 1
    assert st["true_rTr"] <= 4e-12 * st["rTr0"]
    err, err_ref = _rel(pr.p0.f.numpy(), p64), _rel(p32, p64)
    assert err < 1.5 * err_ref + 2e-6, (err, err_ref)
