"""Writes the small parity fixtures of tests/golden/ (run from the repo root:
python tests/golden/make_golden.py).

The reference ships no golden vectors for this path (it has no tests), so the
fixtures are produced by the CPU oracle and each is cross-checked at write
time against an independent computation (scipy's sparse direct solve of the
assembled matrix; min over children for the tag pyramid).  They pin future
changes of the oracle and give the GPU tests fixed inputs/outputs.
"""
import os
import sys

import numpy as np
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from flub_amd import scenes  # noqa: E402
from lapref import assemble  # noqa: E402
from oracle import oracle as O  # noqa: E402


def projection_fixture(name, s, levels=3):
    n = s["n"]
    tag = s["tag"]
    args = (s["u"], s["v"], s["w"], s["ur"], s["vr"], s["wr"], tag, 1.0 / n)
    m64 = O.MGPCG((n,) * 3, levels, real=np.float64, pre_and_post_smoothing=2, bottom_smoothing=6)
    u64, v64, w64, p64, div64, it64, _ = m64.project(*args, rel_tol=1e-26)
    m32 = O.MGPCG((n,) * 3, levels, pre_and_post_smoothing=2, bottom_smoothing=6)
    u32, v32, w32, p32, div32, it32, rtr32 = m32.project(*args, rel_tol=1e-12)
    # independent check of the FP64 solution
    fl = tag == 1
    F = O.divergence_tag(*args[:7], real=np.float64)
    A, _ = assemble(tag)
    y = spla.spsolve(A.tocsc(), F[fl])
    assert np.abs(p64[fl] * n + y).max() < 1e-9 * np.abs(y).max()
    # tag pyramid and divd
    pyr = [tag]
    for _ in range(levels - 1):
        t = pyr[-1]
        c = t.reshape(t.shape[0] // 2, 2, t.shape[1] // 2, 2, t.shape[2] // 2, 2).min(axis=(1, 3, 5))
        assert np.array_equal(O.restrict_tag(t), c)
        pyr.append(c)
    out = dict(tag=tag, u=s["u"], v=s["v"], w=s["w"], ur=s["ur"], vr=s["vr"], wr=s["wr"],
               p64=p64, u64=u64, v64=v64, w64=w64, div32=div32, p32=p32, it32=np.int32(it32),
               levels=np.int32(levels))
    for l, t in enumerate(pyr):
        out[f"tag{l}"] = t.astype(np.int8)
        out[f"divd{l}"] = O.compute_divd(t).astype(np.int8)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    print(f"{name}: fp64 {it64} it, fp32 {it32} it, wrote {name}.npz")


def vcycle_fixture():
    s = scenes.dam_break_spheres(16, 1, n_spheres=4, r_cells=(1, 3))
    tag = s["tag"]
    r = np.random.default_rng(123).normal(size=tag.shape).astype(np.float32)
    r[tag != 1] = 0.0
    m = O.MGPCG(tag.shape, 3, pre_and_post_smoothing=2, bottom_smoothing=6)
    m.set_tags(tag)
    m.put("r", r, 0)
    m.apply_preconditioner("phases")
    np.savez_compressed(os.path.join(HERE, "vcycle_16.npz"), tag=tag, r=r, z=m.get("z", 0),
                        r1=m.get("r", 1), r2=m.get("r", 2))
    print("vcycle_16: wrote vcycle_16.npz")


if __name__ == "__main__":
    projection_fixture("dam_break_16", scenes.dam_break(16, 0))
    projection_fixture("sphere_flow_16", scenes.flow_past_sphere(16, 2))
    vcycle_fixture()
