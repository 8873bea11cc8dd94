"""GPU tests of the mixed-precision solver (FP16/BF16 V-cycle storage inside
FP32 flexible CG with FP64 reductions).  No reference counterpart exists; the
FP64 oracle is the judge: the mixed solve must be as accurate as the FP32
reference algorithm (pressure error against the converged FP64 solution and
post-projection divergence, i.e. the FP64 true residual)."""
import numpy as np
import pytest

from flub_amd import device as wp
from flub_amd import scenes
from flub_amd.multigrid.mgpcg import MGPCG
from flub_amd.project import Projection

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


CASES = [("sphere_flow_64", lambda: scenes.flow_past_sphere(64, 2)),
         ("spheres_64", lambda: scenes.dam_break_spheres(64, 1)),
         ("dam_break_64", lambda: scenes.dam_break(64, 0))]


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
@pytest.mark.parametrize("name,make", CASES)
def test_mixed_projection_accuracy(precision, name, make, oracle):
    s = make()
    n = s["n"]
    pr = Projection((n, n, n), n_mg_levels=4, pre_and_post_smoothing=2, bottom_smoothing=6, precision=precision)
    pr.load(s)
    it = pr.project(max_iters=200, rel_tol=1e-12)
    assert 0 < it < 200
    p = pr.p0.f.numpy()
    u, v, w = pr.velocity()
    o64 = oracle.MGPCG((n, n, n), 4, real=np.float64, pre_and_post_smoothing=2, bottom_smoothing=6)
    _, _, _, p64, _, _, _ = o64.project(s["u"], s["v"], s["w"], s["ur"], s["vr"], s["wr"], s["tag"], 1.0 / n,
                                        rel_tol=1e-24)
    o32 = oracle.MGPCG((n, n, n), 4, pre_and_post_smoothing=2, bottom_smoothing=6)
    u32, v32, w32, p32, _, it32, _ = o32.project(s["u"], s["v"], s["w"], s["ur"], s["vr"], s["wr"], s["tag"],
                                                 1.0 / n, rel_tol=1e-12)
    err, err_ref = _rel(p, p64), _rel(p32, p64)
    assert err < 1.5 * err_ref + 2e-6, (err, err_ref)
    # post-projection divergence (= the true residual of x) no worse than the
    # FP32 reference algorithm's, which stops on its recursive residual
    fl = s["tag"] == 1
    pre = oracle.divergence_tag(s["u"], s["v"], s["w"], s["ur"], s["vr"], s["wr"], s["tag"], real=np.float64)
    post = oracle.divergence_tag(u, v, w, s["ur"], s["vr"], s["wr"], s["tag"], real=np.float64)
    post32 = oracle.divergence_tag(u32, v32, w32, s["ur"], s["vr"], s["wr"], s["tag"], real=np.float64)
    ratio, ratio32 = (np.linalg.norm(x[fl]) / np.linalg.norm(pre[fl]) for x in (post, post32))
    assert ratio < 2.0 * ratio32 + 1e-6, (ratio, ratio32)
    assert it <= 2 * it32 + 4


@pytest.mark.parametrize("precision,tol", [("fp16", 5e-3), ("bf16", 4e-2)])
def test_mixed_vcycle_close_to_fp32(precision, tol):
    s = scenes.dam_break_spheres(64, 1)
    tag = s["tag"]
    r = np.random.default_rng(4).normal(0, 1, tag.shape).astype(np.float32)
    r[tag != 1] = 0.0
    zs = {}
    for prec in ("fp32", precision):
        mg = MGPCG(3, tag.shape, 4, precision=prec)
        mg.pre_and_post_smoothing = 2
        mg.bottom_smoothing = 6
        mg.set_tags(wp.from_numpy(tag))
        mg.r[0].assign(r)
        mg.apply_preconditioner("phases")
        zs[prec] = mg.z[0].numpy()
    assert _rel(zs[precision], zs["fp32"]) < tol


@pytest.mark.parametrize("precision,tol", [("fp16", 5e-3), ("bf16", 4e-2)])
def test_mixed_split_vcycle_close_to_fp32(precision, tol):
    """The V-cycle the mixed solve runs (colour-split level 0: FP16/BF16 r and
    pre-smoothing, FP32 post-smoothing and coarse levels) against the FP32 one."""
    import ctypes
    from flub_amd import _lib
    s = scenes.dam_break_spheres(64, 1)
    tag = s["tag"]
    r = np.random.default_rng(5).normal(0, 1, tag.shape).astype(np.float32)
    r[tag != 1] = 0.0
    zs = {}
    for prec in ("fp32", precision):
        mg = MGPCG(3, tag.shape, 4, precision=prec)
        mg.pre_and_post_smoothing = 2
        mg.bottom_smoothing = 6
        mg.set_tags(wp.from_numpy(tag))
        rd, zd = wp.from_numpy(r), wp.zeros(tag.shape)
        _lib.call("flub_mgpcg_precondition_into", mg.handle, ctypes.c_void_p(rd.ptr), ctypes.c_void_p(zd.ptr), 1)
        zs[prec] = zd.numpy()
    assert np.all(np.isfinite(zs[precision]))
    assert _rel(zs[precision], zs["fp32"]) < tol
