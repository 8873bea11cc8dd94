"""GPU parity of the HIP path against the CPU oracle (through the C-ABI).

Integer work (tags, coarse masks, divd) must be bit-exact; every FP32 kernel
follows the reference's per-cell operation order and must be bitwise equal to
the FP32 oracle on the same inputs; whole solves are compared within stated
tolerances against the FP32 oracle and the FP64 oracle.
"""
import numpy as np
import pytest

from flub_amd import device as wp
from flub_amd import scenes
from flub_amd.const_param import Field, IntField, MACGrid
from flub_amd.divergence import divergence_tag, mask_with_tag
from flub_amd.geometry import tag_dirichlet, tag_neumann
from flub_amd.launch import launch
from flub_amd.multigrid.mgpcg import MGPCG
from flub_amd.project import Projection
from flub_amd.subtract_gradient import subtract_gradient_tag
from flub_amd.utils.util import scale_by_3d

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def random_tags(shape, seed, p=(0.25, 0.6, 0.15)):
    rng = np.random.default_rng(seed)
    return rng.choice(3, size=shape, p=list(p)).astype(np.int32)


SCENE_CASES = [("dam_break", lambda: scenes.dam_break(32, 0)),
               ("spheres", lambda: scenes.dam_break_spheres(32, 1, n_spheres=8, r_cells=(2, 5))),
               ("sphere_flow", lambda: scenes.flow_past_sphere(32, 2))]


@pytest.mark.parametrize("name,make", SCENE_CASES)
def test_tag_classification_bitexact(name, make, oracle):
    s = make()
    phi = Field(f=wp.from_numpy(s["phi"]))
    phr = Field(f=wp.from_numpy(s["phi_rigid"]))
    tag = IntField(f=wp.zeros(s["phi"].shape, int))
    tag_dirichlet(tag, phi)
    tag_neumann(tag, phr)
    ref = oracle.tag_neumann(s["phi_rigid"], oracle.tag_dirichlet(s["phi"]))
    assert np.array_equal(tag.f.numpy(), ref)


@pytest.mark.parametrize("shape,levels,seed", [((16, 16, 16), 4, 0), ((32, 32, 32), 4, 1),
                                               ((64, 32, 16), 3, 2), ((128, 128, 128), 5, 3)])
def test_set_tags_pyramid_and_divd_bitexact(shape, levels, seed, oracle):
    tag = random_tags(shape, seed)
    mg = MGPCG(3, shape, levels)
    mg.set_tags(wp.from_numpy(tag))
    t = tag
    for l in range(levels):
        got = mg.tag[l].numpy().astype(np.int32)
        assert np.array_equal(got, t), f"level {l}"
        assert bits_equal(mg.divd[l].numpy(), oracle.compute_divd(t)), f"divd level {l}"
        if l + 1 < levels:
            t = oracle.restrict_tag(t)


def _mac(s):
    vel = MACGrid(*(Field(f=wp.from_numpy(s[c]), d=d) for d, c in ((1, "u"), (2, "v"), (3, "w"))))
    rig = MACGrid(*(Field(f=wp.from_numpy(s[c]), d=d) for d, c in ((1, "ur"), (2, "vr"), (3, "wr"))))
    return vel, rig


@pytest.mark.parametrize("name,make", SCENE_CASES)
def test_divergence_mask_scale_bitexact(name, make, oracle):
    s = make()
    rng = np.random.default_rng(5)
    for c in ("ur", "vr", "wr"):   # non-zero rigid velocities exercise the solid override
        s[c] = rng.normal(0, 1, s[c].shape).astype(np.float32)
    vel, rig = _mac(s)
    tag = IntField(f=wp.from_numpy(s["tag"]))
    div = Field(f=wp.zeros(s["tag"].shape))
    shape = s["tag"].shape
    launch(divergence_tag, dim=shape, inputs=[vel, div, rig, tag])
    ref = oracle.divergence_tag(s["u"], s["v"], s["w"], s["ur"], s["vr"], s["wr"], s["tag"])
    assert bits_equal(div.f.numpy(), ref)
    launch(mask_with_tag, dim=shape, inputs=[div, tag])
    ref = oracle.mask_with_tag(ref, s["tag"])
    assert bits_equal(div.f.numpy(), ref)
    launch(scale_by_3d, dim=shape, inputs=[div, -1.0 / s["n"]])
    ref = oracle.scale(ref, np.float32(-1.0 / s["n"]))
    assert bits_equal(div.f.numpy(), ref)


@pytest.mark.parametrize("d", [1, 2, 3])
def test_subtract_gradient_bitexact(d, oracle):
    s = scenes.flow_past_sphere(32, 2)
    rng = np.random.default_rng(7)
    p = rng.normal(0, 1, s["tag"].shape).astype(np.float32)
    for c in ("ur", "vr", "wr"):
        s[c] = rng.normal(0, 1, s[c].shape).astype(np.float32)
    vel, rig = _mac(s)
    tag = IntField(f=wp.from_numpy(s["tag"]))
    pf = Field(f=wp.from_numpy(p))
    comp = {1: "u", 2: "v", 3: "w"}[d]
    launch(subtract_gradient_tag, dim=s[comp].shape, inputs=[vel, rig, pf, tag, d])
    got = {1: vel.u, 2: vel.v, 3: vel.w}[d].f.numpy()
    ref = oracle.subtract_gradient_tag(s[comp], s[comp + "r"], p, s["tag"], d)
    assert bits_equal(got, ref)


@pytest.mark.parametrize("shape", [(12, 10, 14), (16, 12, 20)])
def test_divergence_and_gradient_random_bitexact(shape, oracle):
    """Random tags (a fifth SOLID) and velocities on grids whose rows take the
    4-wide kernels (nz % 4 == 0) and the scalar ones (nz = 14)."""
    nx, ny, nz = shape
    rng = np.random.default_rng(sum(shape))
    tag = random_tags(shape, 3, (0.35, 0.45, 0.2))
    sh = {"u": (nx + 1, ny, nz), "v": (nx, ny + 1, nz), "w": (nx, ny, nz + 1)}
    s = {c: rng.normal(0, 1, sh[c[0]]).astype(np.float32) for c in ("u", "v", "w", "ur", "vr", "wr")}
    vel, rig = _mac(s)
    tg = IntField(f=wp.from_numpy(tag))
    div = Field(f=wp.zeros(shape))
    launch(divergence_tag, dim=shape, inputs=[vel, div, rig, tg])
    assert bits_equal(div.f.numpy(), oracle.divergence_tag(s["u"], s["v"], s["w"], s["ur"], s["vr"], s["wr"], tag))
    p = rng.normal(0, 1, shape).astype(np.float32)
    pf = Field(f=wp.from_numpy(p))
    for d, c in ((1, "u"), (2, "v"), (3, "w")):
        launch(subtract_gradient_tag, dim=sh[c], inputs=[vel, rig, pf, tg, d])
        got = {1: vel.u, 2: vel.v, 3: vel.w}[d].f.numpy()
        assert bits_equal(got, oracle.subtract_gradient_tag(s[c], s[c + "r"], p, tag, d)), c


def _loaded_mgpcg(tag, levels, pre_post=2, bottom=6):
    mg = MGPCG(3, tag.shape, levels)
    mg.pre_and_post_smoothing = pre_post
    mg.bottom_smoothing = bottom
    mg.set_tags(wp.from_numpy(tag))
    return mg


@pytest.mark.parametrize("smoother", ["phases", "default"])
def test_compute_Ap_bitexact(smoother, oracle):
    s = scenes.dam_break_spheres(32, 1, n_spheres=8, r_cells=(2, 5))
    tag = s["tag"]
    mg = _loaded_mgpcg(tag, 3)
    p = np.random.default_rng(3).normal(0, 1, tag.shape).astype(np.float32)
    mg.p.assign(p)
    mg.compute_Ap(smoother)
    ref = oracle.compute_Ap(p, tag if smoother == "phases" else None)
    assert bits_equal(mg.Ap.numpy(), ref)
    # the fused FP64 p.Ap lands in `sum`
    assert np.isclose(mg.scalars()[2], float(np.dot(p.astype(np.float64).ravel(), ref.astype(np.float64).ravel())),
                      rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("smoother,shape,levels", [("phases", (32, 32, 32), 3), ("phases", (64, 64, 64), 4),
                                                   ("phases", (64, 32, 32), 4), ("default", (32, 32, 32), 3),
                                                   ("phases", (128, 128, 128), 4), ("phases", (256, 256, 256), 4)])
def test_vcycle_bitexact(smoother, shape, levels, oracle):
    n = shape[0]
    if smoother == "default":
        tag = np.ones(shape, np.int32)
    else:
        s = scenes.dam_break_spheres(n, 1, n_spheres=8, r_cells=(2, 5)) if shape[1] == n else None
        tag = s["tag"] if s is not None else random_tags(shape, 9, (0.1, 0.85, 0.05))
    r = np.random.default_rng(11).normal(0, 1, shape).astype(np.float32)
    r[tag != 1] = 0.0
    mg = _loaded_mgpcg(tag, levels)
    mg.r[0].assign(r)
    orc = oracle.MGPCG(shape, levels, pre_and_post_smoothing=2, bottom_smoothing=6)
    orc.set_tags(tag)
    orc.put("r", r, 0)
    orc.apply_preconditioner(smoother)
    mg.apply_preconditioner(smoother)
    assert bits_equal(mg.z[0].numpy(), orc.get("z", 0))
    for l in range(1, levels):
        assert bits_equal(mg.r[l].numpy(), orc.get("r", l)), f"restricted r level {l}"
    mg.graph_precondition(smoother)          # captured graph replays the same cycle
    mg.graph_precondition(smoother)
    assert bits_equal(mg.z[0].numpy(), orc.get("z", 0))


@pytest.mark.parametrize("shape,levels", [((64, 64, 64), 4), ((32, 48, 64), 4)])
def test_vcycle_blocked_g8_bitexact(shape, levels, oracle, monkeypatch):
    """The optional 8-half-sweep blocked configurations (FLUB_TB_G8) on the
    coarse levels of the reference-layout V-cycle: bitwise the oracle."""
    monkeypatch.setenv("FLUB_TB_G8", "1")
    tag = random_tags(shape, 6, (0.1, 0.85, 0.05))
    r = np.random.default_rng(14).normal(0, 1, shape).astype(np.float32)
    r[tag != 1] = 0.0
    mg = _loaded_mgpcg(tag, levels)
    mg.r[0].assign(r)
    orc = oracle.MGPCG(shape, levels, pre_and_post_smoothing=2, bottom_smoothing=6)
    orc.set_tags(tag)
    orc.put("r", r, 0)
    orc.apply_preconditioner("phases")
    mg.apply_preconditioner("phases")
    assert bits_equal(mg.z[0].numpy(), orc.get("z", 0))


@pytest.mark.parametrize("shape,levels", [((64, 64, 64), 4), ((128, 128, 128), 4), ((64, 32, 48), 3),
                                          ((256, 256, 256), 4)])
def test_split_vcycle_bitexact(shape, levels, oracle):
    """The V-cycle the fused FP32 solve runs (level 0 in the colour-split
    layout, blocked coarse levels), reached through precondition_into."""
    import ctypes
    from flub_amd import _lib
    n = shape[0]
    s = scenes.dam_break_spheres(n, 1, n_spheres=8, r_cells=(2, 5)) if len(set(shape)) == 1 else None
    tag = s["tag"] if s is not None else random_tags(shape, 4, (0.1, 0.85, 0.05))
    r = np.random.default_rng(12).normal(0, 1, shape).astype(np.float32)
    r[tag != 1] = 0.0
    mg = _loaded_mgpcg(tag, levels)
    rd, zd = wp.from_numpy(r), wp.zeros(shape)
    _lib.call("flub_mgpcg_precondition_into", mg.handle, ctypes.c_void_p(rd.ptr), ctypes.c_void_p(zd.ptr), 1)
    orc = oracle.MGPCG(shape, levels, pre_and_post_smoothing=2, bottom_smoothing=6)
    orc.set_tags(tag)
    orc.put("r", r, 0)
    orc.apply_preconditioner("phases")
    assert bits_equal(zd.numpy(), orc.get("z", 0))


@pytest.mark.parametrize("variant", [None, "FLUB_SPLIT_GRAPH", "FLUB_NO_PAIR"])
@pytest.mark.parametrize("shape,levels", [((64, 64, 64), 4), ((64, 32, 48), 3), ((32, 48, 64), 4)])
def test_split_coarse_vcycle_bitexact(shape, levels, variant, oracle, monkeypatch):
    """The split V-cycle with every level but the bottom colour-split (the
    layout grids of 2^20 cells and more use on their coarse levels), forced
    down to small grids: bitwise the reference V-cycle, launched kernel by
    kernel with the coarse half-sweeps in pairs (levels whose shape allows it;
    (64, 32, 48) has none), replayed from its captured graph, and with one
    launch per half-sweep."""
    import ctypes
    from flub_amd import _lib
    monkeypatch.setenv("FLUB_SPLIT_MIN_CELLS", "1")
    if variant:
        monkeypatch.setenv(variant, "1")
    tag = random_tags(shape, 5, (0.1, 0.85, 0.05))
    r = np.random.default_rng(13).normal(0, 1, shape).astype(np.float32)
    r[tag != 1] = 0.0
    mg = _loaded_mgpcg(tag, levels)
    rd, zd = wp.from_numpy(r), wp.zeros(shape)
    for _ in range(2):   # with the graph, the second call replays it
        _lib.call("flub_mgpcg_precondition_into", mg.handle, ctypes.c_void_p(rd.ptr), ctypes.c_void_p(zd.ptr), 1)
    orc = oracle.MGPCG(shape, levels, pre_and_post_smoothing=2, bottom_smoothing=6)
    orc.set_tags(tag)
    orc.put("r", r, 0)
    orc.apply_preconditioner("phases")
    assert bits_equal(zd.numpy(), orc.get("z", 0))


@pytest.mark.parametrize("variant", ["FLUB_SPLIT_GRAPH", "FLUB_RTR_COPY", "FLUB_NO_PAIR", "FLUB_MX_SRC32",
                                     "FLUB_MX_NO_FUSE0"])
@pytest.mark.parametrize("precision", ["fp32", "fp16", "bf16"])
def test_split_launch_variants_bitwise(precision, variant, monkeypatch):
    """How the fused solves queue their work changes no value: replaying the
    captured split V-cycle instead of launching it kernel by kernel, copying
    the residual norm to the host instead of storing it there from the finish
    kernel, one launch per coarse half-sweep instead of pairs, or (BF16) the
    first sweeps converting the scaled FP32 r instead of reading the BF16 r
    the r update stored at scale 1, or (BF16) the first half-sweep launched on
    its own instead of run by the r update, gives bitwise the same pressure in
    the same iterations (coarse levels split as on large grids)."""
    s = scenes.dam_break_spheres(64, 1)
    monkeypatch.setenv("FLUB_SPLIT_MIN_CELLS", "1")
    if variant in ("FLUB_NO_PAIR", "FLUB_MX_SRC32"):
        # the BF16 level-1 pre-smoothing needs the pairs and the BF16 r update
        monkeypatch.setenv("FLUB_MX_NO_L1", "1")
    res = []
    for on in (False, True):
        if on:
            monkeypatch.setenv(variant, "1")
        pr = Projection((64, 64, 64), n_mg_levels=4, pre_and_post_smoothing=2, bottom_smoothing=6,
                        precision=precision)
        pr.load(s)
        it = pr.project(max_iters=-1, rel_tol=1e-12)
        res.append((it, pr.p0.f.numpy(), pr.rTr))
    (i0, p0, r0), (i1, p1, r1) = res
    assert i0 == i1 and r0 == r1
    assert bits_equal(p0, p1)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_split_graph_across_solves(precision, monkeypatch):
    """Captured split V-cycles reused across solves on one handle (a capped
    solve, then full solves of two scenes): the BF16 cycle's first sweeps
    depend on what the r update stored, so a captured cycle is replayed only
    in the state it was captured in; the pressures equal the eager path's
    bitwise."""
    monkeypatch.setenv("FLUB_SPLIT_MIN_CELLS", "1")
    scs = [scenes.dam_break_spheres(64, 1), scenes.dam_break(64, 3), scenes.dam_break_spheres(64, 2)]
    caps = [3, -1, -1]
    res = []
    for graph in (False, True):
        if graph:
            monkeypatch.setenv("FLUB_SPLIT_GRAPH", "1")
        pr = Projection((64, 64, 64), n_mg_levels=4, pre_and_post_smoothing=2, bottom_smoothing=6,
                        precision=precision)
        out = []
        for sc, cap in zip(scs, caps):
            pr.load(sc)
            it = pr.project(max_iters=cap, rel_tol=1e-12)
            out.append((it, pr.p0.f.numpy().copy()))
        res.append(out)
    for (i0, p0), (i1, p1) in zip(*res):
        assert i0 == i1
        assert bits_equal(p0, p1)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_split_solve_iteration_cap(precision, monkeypatch):
    """A fused solve stopped by max_iters returns that count; its pressure and
    the r.r it reports (summed on the host from the r update's partials) are
    bitwise those of the path with a finish kernel and a copy."""
    s = scenes.dam_break_spheres(64, 1)
    monkeypatch.setenv("FLUB_SPLIT_MIN_CELLS", "1")
    res = []
    for copy in (False, True):
        if copy:
            monkeypatch.setenv("FLUB_RTR_COPY", "1")
        pr = Projection((64, 64, 64), n_mg_levels=4, pre_and_post_smoothing=2, bottom_smoothing=6,
                        precision=precision)
        pr.load(s)
        it = pr.project(max_iters=3, rel_tol=1e-12)
        res.append((it, pr.p0.f.numpy(), pr.rTr))
    (i0, p0, r0), (i1, p1, r1) = res
    assert i0 == i1 == 3
    assert r0 == r1 and np.isfinite(r0) and r0 > 0.0
    assert bits_equal(p0, p1)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_split_solve_nonfinite(precision):
    """A non-finite right-hand side stops the fused solve with the non-finite
    status instead of iterating on NaN."""
    from flub_amd import _lib
    s = scenes.dam_break_spheres(64, 1)
    pr = Projection((64, 64, 64), n_mg_levels=4, pre_and_post_smoothing=2, bottom_smoothing=6,
                    precision=precision)
    pr.load(s)
    u = np.full_like(np.asarray(s["u"], np.float32), np.nan)
    pr.set_velocity(u, s["v"], s["w"])
    with pytest.raises(_lib.FlubError) as e:
        pr.project(max_iters=-1, rel_tol=1e-12)
    assert e.value.code == _lib.ENONFINITE


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_split_coarse_solve_bitwise(precision, monkeypatch):
    """Colour-split coarse levels change where values are stored, not the
    arithmetic: the solve gives bitwise the same pressure in the same
    iterations with and without them (BF16: with level 1 kept in FP32)."""
    s = scenes.dam_break_spheres(64, 1)
    monkeypatch.setenv("FLUB_MX_NO_L1", "1")
    res = []
    for cells in ("1", str(1 << 30)):
        monkeypatch.setenv("FLUB_SPLIT_MIN_CELLS", cells)
        pr = Projection((64, 64, 64), n_mg_levels=4, pre_and_post_smoothing=2, bottom_smoothing=6,
                        precision=precision)
        pr.load(s)
        it = pr.project(max_iters=-1, rel_tol=1e-12)
        res.append((it, pr.p0.f.numpy()))
    (i0, p0), (i1, p1) = res
    assert i0 == i1
    assert bits_equal(p0, p1)


@pytest.mark.parametrize("name,make,cells", [("spheres_64", lambda: scenes.dam_break_spheres(64, 1), "1"),
                                             ("sphere_flow_256", lambda: scenes.flow_past_sphere(256, 2), None)])
def test_bf16_level1_presmoothing(name, make, cells, monkeypatch):
    """The BF16 solve pre-smooths a split level 1 in BF16 (r and z; the
    post-smoothing stays FP32): against the same solve with level 1 in FP32,
    the same iterations (at most one more) and the same FP64 true residual
    test passed; the pressures agree to the solves' accuracy."""
    s = make()
    n = s["n"]
    if cells:
        monkeypatch.setenv("FLUB_SPLIT_MIN_CELLS", cells)
    res = []
    for off in (False, True):
        if off:
            monkeypatch.setenv("FLUB_MX_NO_L1", "1")
        pr = Projection((n, n, n), n_mg_levels=4, pre_and_post_smoothing=2, bottom_smoothing=6, precision="bf16")
        pr.load(s)
        it = pr.project(max_iters=-1, rel_tol=1e-12)
        div = pr.div.f.numpy()
        b = np.where(s["tag"] == 1, div * np.float32(-pr.h), np.float32(0)).astype(np.float64)
        res.append((it, pr.p0.f.numpy().astype(np.float64), pr.rTr, float(np.sum(b * b))))
    (i1, p1, r1, bb), (i0, p0, r0, _) = res
    rel = float(np.linalg.norm(p1 - p0) / np.linalg.norm(p0))
    print(f"{name}: BF16 level 1 {i1} iterations, FP32 level 1 {i0}; true rel. residual "
          f"{np.sqrt(r1 / bb):.3e} / {np.sqrt(r0 / bb):.3e}; pressures differ by {rel:.2e}")
    assert i0 <= i1 <= i0 + 1
    assert r1 < 1e-12 * bb and r0 < 1e-12 * bb
    assert rel < 1e-5


def test_split_solve_matches_reference_layout(monkeypatch, oracle):
    """The fused solve on the split level 0 (FP64 x) and on the reference
    layout (the reference's FP32 x) converge to the same solution: the split
    one at least as close to the converged FP64 solution as the reference
    algorithm, in at most one iteration more (its FP64 true-residual check)."""
    s = scenes.dam_break_spheres(64, 1)
    res = []
    for env in (None, "1"):
        if env:
            monkeypatch.setenv("FLUB_NO_SPLIT", env)
        pr = Projection((64, 64, 64), n_mg_levels=4, pre_and_post_smoothing=2, bottom_smoothing=6)
        pr.load(s)
        it = pr.project(max_iters=-1, rel_tol=1e-12)
        res.append((it, pr.p0.f.numpy(), pr.rTr))
    (i0, p0, r0), (i1, p1, r1) = res
    o64 = oracle.MGPCG((64,) * 3, 4, real=np.float64, pre_and_post_smoothing=2, bottom_smoothing=6)
    _, _, _, p64, _, _, _ = o64.project(s["u"], s["v"], s["w"], s["ur"], s["vr"], s["wr"], s["tag"], 1.0 / 64,
                                        rel_tol=1e-24)
    assert 0 <= i0 - i1 <= 1
    assert _rel(p0, p1) < 2e-6
    assert _rel(p0, p64) <= _rel(p1, p64) * 1.01 + 1e-7
    assert r0 < 1.0 and r1 < 1.0


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


@pytest.mark.parametrize("name,make,levels", [("dam_break_64", lambda: scenes.dam_break(64, 0), 4),
                                              ("spheres_64", lambda: scenes.dam_break_spheres(64, 1), 4),
                                              ("sphere_flow_64", lambda: scenes.flow_past_sphere(64, 2), 4)])
def test_projection_matches_oracle(name, make, levels, oracle):
    s = make()
    n = s["n"]
    pr = Projection((n, n, n), n_mg_levels=levels, pre_and_post_smoothing=2, bottom_smoothing=6)
    pr.load(s)
    it = pr.project(max_iters=-1, rel_tol=1e-12)
    p_gpu = pr.p0.f.numpy()
    u, v, w = pr.velocity()
    o32 = oracle.MGPCG((n, n, n), levels, pre_and_post_smoothing=2, bottom_smoothing=6)
    ru, rv, rw, rp, rdiv, rit, rrtr = o32.project(s["u"], s["v"], s["w"], s["ur"], s["vr"], s["wr"], s["tag"],
                                                   1.0 / n, rel_tol=1e-12)
    o64 = oracle.MGPCG((n, n, n), levels, real=np.float64, pre_and_post_smoothing=2, bottom_smoothing=6)
    _, _, _, p64, _, _, _ = o64.project(s["u"], s["v"], s["w"], s["ur"], s["vr"], s["wr"], s["tag"], 1.0 / n,
                                        rel_tol=1e-24)
    assert bits_equal(pr.div.f.numpy(), rdiv)                 # fused divergence == reference sequence
    assert abs(it - rit) <= 1
    assert pr.rTr is not None and pr.rTr < 1.0
    assert _rel(p_gpu, rp) < 1e-5                             # FP32 GPU vs FP32 oracle
    # vs the converged FP64 solution (residual 1e-12): a 1e-6 residual bounds the
    # error only up to the conditioning, so the GPU must be as accurate as the
    # FP32 reference algorithm itself (plus rounding slack)
    err_ref = _rel(rp, p64)
    assert _rel(p_gpu, p64) < 1.5 * err_ref + 1e-6, (_rel(p_gpu, p64), err_ref)
    assert _rel(u, ru) < 1e-5 and _rel(v, rv) < 1e-5 and _rel(w, rw) < 1e-5
    # post-projection divergence on FLUID cells equals the solver's residual (small)
    post = oracle.divergence_tag(u, v, w, s["ur"], s["vr"], s["wr"], s["tag"], real=np.float64)
    pre = oracle.divergence_tag(s["u"], s["v"], s["w"], s["ur"], s["vr"], s["wr"], s["tag"], real=np.float64)
    fl = s["tag"] == 1
    assert np.linalg.norm(post[fl]) < 2e-5 * np.linalg.norm(pre[fl])


def test_stepwise_equals_fused(oracle):
    """The fused projection and the launch-by-launch sequence solve the same
    way (bitwise the same pressure in the same iterations).  The fused one
    subtracts the gradient of the solve's FP64 x instead of the FP32 pressure
    chain: its velocities differ by that rounding and are at least as
    divergence-free."""
    s = scenes.dam_break_spheres(64, 1)
    a = Projection((64,) * 3, 4, 2, 6)
    a.load(s)
    a.project()
    b = Projection((64,) * 3, 4, 2, 6)
    b.load(s)
    b.project_stepwise()
    assert a.iterations == b.iterations
    assert bits_equal(a.p0.f.numpy(), b.p0.f.numpy())
    fl = s["tag"] == 1
    post = []
    for pr in (a, b):
        u, v, w = pr.velocity()
        post.append(np.linalg.norm(oracle.divergence_tag(u, v, w, s["ur"], s["vr"], s["wr"], s["tag"],
                                                         real=np.float64)[fl]))
    for x, y in zip(a.velocity(), b.velocity()):
        assert _rel(x, y) < 1e-6
    assert post[0] <= post[1] * 1.001


def test_half_fluid_cos_problem(oracle):
    """The reference's MGPCG_Test problem (mgpcg.py:501-558) at 64^3, 4 levels."""
    tag, r = scenes.half_fluid_cos(64)
    mg = _loaded_mgpcg(tag, 4, pre_post=6, bottom=6)
    mg.init(wp.from_numpy(r), 1.0)
    mg.mask_interior()
    it = mg.solve(400, rel_tol=1e-12, abs_tol=0.0, smoother="phases")
    o = oracle.MGPCG(tag.shape, 4, real=np.float64)
    o.set_tags(tag)
    o.init(r.astype(np.float64), 1.0)
    o.mask_interior()
    oit, _ = o.solve(400, eps=0.0, rel_tol=1e-24, abs_tol=0.0, smoother="phases")
    assert it < 400
    assert _rel(mg.x.numpy(), o.get("x")) < 1e-5


def test_residual_norm_fp64(oracle):
    """flub_mgpcg_residual_norm against the FP64 oracle operator."""
    import ctypes
    from flub_amd import _lib
    tag = random_tags((32, 32, 32), 21, (0.2, 0.7, 0.1))
    rng = np.random.default_rng(22)
    x = rng.normal(0, 1, tag.shape).astype(np.float32)
    b = rng.normal(0, 1, tag.shape).astype(np.float32)
    mg = _loaded_mgpcg(tag, 3)
    xd, bd = wp.from_numpy(x), wp.from_numpy(b)
    rr = ctypes.c_double()
    _lib.call("flub_mgpcg_residual_norm", mg.handle, xd.ptr, bd.ptr, ctypes.byref(rr))
    ndt = oracle.compute_Ap(x.astype(np.float64), tag, real=np.float64)   # neighbor_dot_tag(x), 0 off FLUID
    res = np.where(tag == 1, b.astype(np.float64) - ndt, 0.0)
    assert np.isclose(rr.value, float(np.sum(res * res)), rtol=1e-12)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("shape", [(64, 64, 64), (48, 32, 48)])
def test_fused_gradient_equals_per_axis(shape, precision, monkeypatch):
    """The fused projection's gradient subtraction in one pass over the three
    axes from the colour-split FP64 x equals the per-axis passes over its
    reference-layout copy (FLUB_GRAD_PER_AXIS=1) bitwise, boundary faces
    included."""
    n = shape[0]
    s = scenes.dam_break_spheres(n, 1) if len(set(shape)) == 1 else None
    if s is None:
        rng = np.random.default_rng(41)
        tag = rng.choice(3, size=shape, p=[0.15, 0.75, 0.10]).astype(np.int32)
        nx, ny, nz = shape
        s = {"tag": tag, "u": rng.normal(0, 1, (nx + 1, ny, nz)).astype(np.float32),
             "v": rng.normal(0, 1, (nx, ny + 1, nz)).astype(np.float32),
             "w": rng.normal(0, 1, (nx, ny, nz + 1)).astype(np.float32),
             "ur": rng.normal(0, 1, (nx + 1, ny, nz)).astype(np.float32),
             "vr": rng.normal(0, 1, (nx, ny + 1, nz)).astype(np.float32),
             "wr": rng.normal(0, 1, (nx, ny, nz + 1)).astype(np.float32)}
    res = []
    for env in (None, "1"):
        if env:
            monkeypatch.setenv("FLUB_GRAD_PER_AXIS", env)
        pr = Projection(shape, n_mg_levels=3, pre_and_post_smoothing=2, bottom_smoothing=6, precision=precision)
        pr.load(s)
        pr.project(max_iters=-1, rel_tol=1e-12)
        res.append([a.copy() for a in pr.velocity()] + [pr.p0.f.numpy()])
    for a, b in zip(*res):
        assert bits_equal(a, b)
