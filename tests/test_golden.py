"""Committed fixtures (tests/golden/, written by make_golden.py and checked
there against scipy's direct solve): the oracle must keep reproducing them
(CPU), and the HIP path must match them (GPU: tags/divd/divergence/V-cycle
bit-exact, pressure within tolerance)."""
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PROJ = ["dam_break_16", "sphere_flow_16"]


def load(name):
    with np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def _inputs(g):
    return g["u"], g["v"], g["w"], g["ur"], g["vr"], g["wr"]


@pytest.mark.parametrize("name", PROJ)
def test_oracle_reproduces_projection_fixture(name, oracle):
    g = load(name)
    n, L = g["tag"].shape[0], int(g["levels"])
    m64 = oracle.MGPCG((n,) * 3, L, real=np.float64, pre_and_post_smoothing=2, bottom_smoothing=6)
    u, v, w, p, _, _, _ = m64.project(*_inputs(g), g["tag"], 1.0 / n, rel_tol=1e-26)
    assert np.allclose(p, g["p64"], rtol=0, atol=1e-12 * np.abs(g["p64"]).max())
    m32 = oracle.MGPCG((n,) * 3, L, pre_and_post_smoothing=2, bottom_smoothing=6)
    _, _, _, p32, div32, it32, _ = m32.project(*_inputs(g), g["tag"], 1.0 / n, rel_tol=1e-12)
    assert np.array_equal(div32.view(np.uint32), g["div32"].view(np.uint32))
    assert it32 == int(g["it32"]) and np.array_equal(p32.view(np.uint32), g["p32"].view(np.uint32))
    t = g["tag"]
    for l in range(L):
        assert np.array_equal(t, g[f"tag{l}"])
        assert np.array_equal(oracle.compute_divd(t), g[f"divd{l}"])
        t = oracle.restrict_tag(t) if l + 1 < L else t


@pytest.mark.gpu
@pytest.mark.parametrize("name", PROJ)
def test_hip_matches_projection_fixture(name):
    from flub_amd import device as wp
    from flub_amd.multigrid.mgpcg import MGPCG
    from flub_amd.project import Projection

    g = load(name)
    n, L = g["tag"].shape[0], int(g["levels"])
    mg = MGPCG(3, (n,) * 3, L)
    mg.set_tags(wp.from_numpy(g["tag"]))
    for l in range(L):
        assert np.array_equal(mg.tag[l].numpy().astype(np.int8), g[f"tag{l}"])
        assert np.array_equal(mg.divd[l].numpy().astype(np.int8), g[f"divd{l}"])
    pr = Projection((n,) * 3, n_mg_levels=L, pre_and_post_smoothing=2, bottom_smoothing=6)
    pr.load(dict(u=g["u"], v=g["v"], w=g["w"], ur=g["ur"], vr=g["vr"], wr=g["wr"], tag=g["tag"]))
    it = pr.project(rel_tol=1e-12)
    assert np.array_equal(pr.div.f.numpy().view(np.uint32), g["div32"].view(np.uint32))
    assert abs(it - int(g["it32"])) <= 1
    p = pr.p0.f.numpy()
    assert np.linalg.norm(p - g["p32"]) / np.linalg.norm(g["p32"]) < 1e-5
    err_ref = np.linalg.norm(g["p32"] - g["p64"]) / np.linalg.norm(g["p64"])
    assert np.linalg.norm(p - g["p64"]) / np.linalg.norm(g["p64"]) < 1.5 * err_ref + 1e-6


@pytest.mark.gpu
def test_hip_matches_vcycle_fixture():
    from flub_amd import device as wp
    from flub_amd.multigrid.mgpcg import MGPCG

    g = load("vcycle_16")
    mg = MGPCG(3, g["tag"].shape, 3)
    mg.pre_and_post_smoothing = 2
    mg.bottom_smoothing = 6
    mg.set_tags(wp.from_numpy(g["tag"]))
    mg.r[0].assign(g["r"])
    mg.apply_preconditioner("phases")
    assert np.array_equal(mg.z[0].numpy().view(np.uint32), g["z"].view(np.uint32))
    assert np.array_equal(mg.r[1].numpy().view(np.uint32), g["r1"].view(np.uint32))
    assert np.array_equal(mg.r[2].numpy().view(np.uint32), g["r2"].view(np.uint32))
