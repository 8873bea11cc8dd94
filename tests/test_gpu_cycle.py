"""The V-cycle's coarsest levels in one workgroup (csrc/tail.hip) and the cycle
variant without a reference counterpart (coarse-grid correction scale, sweep
cap, more levels), against the FP32 oracle through the C-ABI:

* the split V-cycle the fused solves run (flub_mgpcg_precondition_into) with
  the one-workgroup tail equals the oracle's V-cycle bitwise, and equals the
  same cycle launched level by level (FLUB_NO_TAIL=1), for the reference cycle
  and for the variant (scale 1/2 or 0.6, cap 2-3, down to 2^3-4^3 cells);
* projections with the variant converge to the converged FP64 oracle
  solution (relative pressure error 1e-5) under the FP64 true-residual stop,
  in FP32 and BF16, in fewer iterations than the reference cycle.
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


def _rel(a, b):
    return float(np.linalg.norm((np.asarray(a, np.float64) - b).ravel()) / np.linalg.norm(np.ravel(b)))


def _tags(shape, seed):
    n = shape[0]
    if len(set(shape)) == 1:
        return scenes.dam_break_spheres(n, seed, n_spheres=8, r_cells=(2, 5))["tag"]
    rng = np.random.default_rng(seed)
    return rng.choice(3, size=shape, p=[0.1, 0.85, 0.05]).astype(np.int32)


def _precondition(tag, r, levels, cs, cap):
    mg = MGPCG(3, tag.shape, levels)
    mg.pre_and_post_smoothing, mg.bottom_smoothing = 2, 6
    if cs != 1.0:
        mg.coarse_scale = cs
    if cap:
        mg.sweep_cap = cap
    mg.set_tags(wp.from_numpy(tag))
    rd, zd = wp.from_numpy(r), wp.zeros(tag.shape)
    _lib.call("flub_mgpcg_precondition_into", mg.handle, ctypes.c_void_p(rd.ptr), ctypes.c_void_p(zd.ptr), 1)
    return zd.numpy()


@pytest.mark.parametrize("shape,levels,cs,cap", [((64, 64, 64), 4, 1.0, 0), ((64, 64, 64), 5, 1.0, 0),
                                                 ((128, 128, 128), 6, 0.5, 3), ((128, 64, 96), 5, 0.6, 2),
                                                 ((256, 256, 256), 7, 0.5, 3), ((128, 128, 128), 7, 0.5, 2)])
def test_tail_vcycle_bitexact(shape, levels, cs, cap, oracle, monkeypatch):
    tag = _tags(shape, 3)
    r = np.random.default_rng(17).normal(0, 1, shape).astype(np.float32)
    r[tag != 1] = 0.0
    z_tail = _precondition(tag, r, levels, cs, cap)
    monkeypatch.setenv("FLUB_NO_TAIL", "1")
    z_launch = _precondition(tag, r, levels, cs, cap)
    orc = oracle.MGPCG(shape, levels, pre_and_post_smoothing=2, bottom_smoothing=6)
    orc.set_tags(tag)
    orc.set_cycle(cs, cap)
    orc.put("r", r, 0)
    orc.apply_preconditioner("phases")
    assert bits_equal(z_tail, z_launch)
    assert bits_equal(z_tail, orc.get("z", 0))


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("name,make", [("sphere_flow_128", lambda: scenes.flow_past_sphere(128, 2)),
                                       ("spheres_128", lambda: scenes.dam_break_spheres(128, 1))])
def test_variant_cycle_projection(name, make, precision, oracle):
    s = make()
    n = s["n"]
    o64 = oracle.MGPCG((n,) * 3, 4, real=np.float64, pre_and_post_smoothing=2, bottom_smoothing=6)
    _, _, _, p64, _, _, _ = o64.project(s["u"], s["v"], s["w"], s["ur"], s["vr"], s["wr"], s["tag"], 1.0 / n,
                                        rel_tol=1e-24)
    del o64
    res = {}
    for cycle in ("reference", "variant"):
        kw = dict(coarse_scale=0.5, sweep_cap=3) if cycle == "variant" else {}
        levels = 6 if cycle == "variant" else 4
        pr = Projection((n,) * 3, n_mg_levels=levels, pre_and_post_smoothing=2, bottom_smoothing=6,
                        precision=precision, **kw)
        pr.load(s)
        it = pr.project(max_iters=-1, rel_tol=1e-12)
        res[cycle] = (it, _rel(pr.p0.f.numpy(), p64))
    print(f"{name} {precision}: reference cycle {res['reference'][0]} iterations (error {res['reference'][1]:.2e}), "
          f"variant {res['variant'][0]} (error {res['variant'][1]:.2e})")
    assert res["variant"][1] < 1e-5
    assert res["variant"][0] < res["reference"][0]


def _isolated_pocket_tags(n=64):
    """SOLID everywhere but a FLUID block (i < n/2) with an AIR face (i = 0),
    a 2x2x2 FLUID pocket whose level-1 cell P = (20, 20, 20) has only SOLID
    face neighbours (divd 0 on level 1), and a line of FLUID level-1 cells
    (16..21, 21, 21) that ties P's level-2 parent to the block."""
    tag = np.full((n, n, n), 2, np.int32)
    tag[: n // 2] = 1
    tag[0] = 0
    tag[40:42, 40:42, 40:42] = 1                      # the pocket: level-1 cell (20, 20, 20)
    tag[32:44, 42:44, 42:44] = 1                      # level-1 cells (16..21, 21, 21)
    return tag


def test_isolated_coarse_fluid_cell_bitexact(oracle):
    """A FLUID cell with no AIR/FLUID neighbour has divd 0 and the reference
    sets its z to 0 on every sweep (mgpcg.py:80-83), also after a
    prolongation added its parent's value; the children of that cell then
    receive 0.  The V-cycle paths (split, reference layout, blocked and tail
    levels) follow it bitwise."""
    tag = _isolated_pocket_tags()
    t1 = oracle.restrict_tag(tag)
    assert t1[20, 20, 20] == 1 and oracle.compute_divd(t1)[20, 20, 20] == 0
    assert oracle.restrict_tag(t1)[10, 10, 10] == 1
    r = np.random.default_rng(23).normal(0, 1, tag.shape).astype(np.float32)
    r[tag != 1] = 0.0
    orc = oracle.MGPCG(tag.shape, 4, pre_and_post_smoothing=2, bottom_smoothing=6)
    orc.set_tags(tag)
    orc.put("r", r, 0)
    orc.apply_preconditioner("phases")
    want = orc.get("z", 0)
    assert orc.get("z", 2)[10, 10, 10] != 0.0      # the parent's value the prolongation passes down
    assert bits_equal(_precondition(tag, r, 4, 1.0, 0), want)
    mg = MGPCG(3, tag.shape, 4)
    mg.pre_and_post_smoothing, mg.bottom_smoothing = 2, 6
    mg.set_tags(wp.from_numpy(tag))
    mg.r[0].assign(r)
    mg.apply_preconditioner("phases")
    assert bits_equal(mg.z[0].numpy(), want)


@pytest.mark.parametrize("precision", ["fp32", "fp16", "bf16"])
@pytest.mark.parametrize("shape", [(64, 64, 64), (128, 64, 96)])
def test_tuned_preset_projection(shape, precision, oracle):
    """The preset the bench's tuned legs use (flub_amd.project.TUNED_CYCLE,
    tuned_levels: levels down to 4 cells on the shortest axis) on a cubic and a
    non-cubic grid (the latter's coarsest levels are not powers of two and run
    launch by launch): pressure within 1e-5 of the converged FP64 oracle, in
    every storage precision, stopping on the FP64 true residual -- or, where the
    1e-6 residual leaves a larger error in the lowest modes, as close as the
    reference cycle's solve of the same scene."""
    from flub_amd.project import TUNED_CYCLE, tuned_levels
    rng = np.random.default_rng(51)
    nx, ny, nz = shape
    tag = np.ones(shape, np.int32)
    tag[:, ny // 2:, :] = 0                       # a free surface
    c = np.array(shape) * np.array([0.3, 0.25, 0.6])
    i, j, k = np.meshgrid(*(np.arange(s) for s in shape), indexing="ij")
    tag[(i - c[0]) ** 2 + (j - c[1]) ** 2 + (k - c[2]) ** 2 < (min(shape) / 6) ** 2] = 2
    u = rng.normal(0, 1, (nx + 1, ny, nz)).astype(np.float32)
    v = rng.normal(0, 1, (nx, ny + 1, nz)).astype(np.float32)
    w = rng.normal(0, 1, (nx, ny, nz + 1)).astype(np.float32)
    zr = [np.zeros_like(a) for a in (u, v, w)]
    s = {"tag": tag, "u": u, "v": v, "w": w, "ur": zr[0], "vr": zr[1], "wr": zr[2]}
    levels = tuned_levels(shape)
    o64 = oracle.MGPCG(shape, 4, real=np.float64, pre_and_post_smoothing=2, bottom_smoothing=6)
    _, _, _, p64, _, _, _ = o64.project(u, v, w, *zr, tag, 1.0 / nx, rel_tol=1e-24)
    errs = {}
    for cycle in ("tuned", "reference"):
        kw = dict(n_mg_levels=levels, **TUNED_CYCLE) if cycle == "tuned" else dict(n_mg_levels=3)
        pr = Projection(shape, pre_and_post_smoothing=2, bottom_smoothing=6, precision=precision, **kw)
        pr.load(s)
        it = pr.project(max_iters=-1, rel_tol=1e-12)
        errs[cycle] = (it, _rel(pr.p0.f.numpy(), p64))
    print(f"{shape} {precision}: tuned {levels} levels, {errs['tuned'][0]} iterations, pressure error vs FP64 "
          f"{errs['tuned'][1]:.2e}; reference cycle {errs['reference'][0]} iterations, {errs['reference'][1]:.2e}")
    assert errs["tuned"][1] < max(1e-5, 2.0 * errs["reference"][1])
    assert errs["tuned"][0] <= errs["reference"][0]
