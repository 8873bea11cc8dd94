"""Benchmark: ms per pressure projection to 1e-6 relative residual on one MI355X.

Workload (default): flow past a sphere on a 256^3 MAC grid -- solid sphere of
radius 0.15 at the centre, inflow u = 1 through a SOLID ghost layer at x-min,
outflow through an AIR layer at x-max, u = 1 + N(0, 0.1), v, w = N(0, 0.1)
(seed 2).  One step = one full projection (divergence, pressure solve to
||r|| / ||r0|| < 1e-6, gradient subtraction) with the inputs resident in HBM;
the velocity inputs are restored by a device copy outside the timed region
before every step.

Solvers, each timed over the same steps: "fp32" (the reference MGPCG:
4-level V-cycle, 2 RB-GS pre/post sweeps with the reference's doubling per
level, 6 bottom sweeps), "fp16" / "bf16" (the same V-cycle with FP16/BF16
storage inside FP32 flexible CG with FP64 reductions), "fp32-tuned" /
"bf16-tuned" (the tuned V-cycle, no reference counterpart: levels down to
4^3, i.e. 7 at 256^3, the coarse-grid correction scaled by 1/2, at most 3
red-black iterations per level, the smallest levels in one workgroup; the
same smoother, transfers and stop rule) and "learned" (U-Net
preconditioner in flexible PCG with V-cycle fallback, weights
flub_amd/learned/weights/unet_128.npz).  Every solver keeps x in FP64 and
stops when the FP64 true residual ||b - A x|| of that x is below 1e-6 ||b||
(the reference's test, rel_tol = 1e-12 on squared norms): one accuracy rule
for all legs.  "bf16" also keeps level 1's pre-smoothing in BF16.  `value` is the fastest solver's ms per projection; every
solver's figures are in "solvers", with its FP64 true residual, the true
residual of the FP32 pressure it hands back, and (the FP64 judge) its
relative pressure error against the converged FP64 CPU oracle projection, the
energy-norm error of its FP64 x, and the post-projection divergence of its
velocity field.  "learned_heldout":
the learned solver against plain MGPCG and V-cycle FPCG on held-out seeds at
128^3 and 256^3.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256]
                    [--solvers fp32 fp16 bf16 fp32-tuned bf16-tuned learned]
                    [--no-cpu] [--no-judge] [--no-heldout] [--no-configs]

"configs" (after the headline, unless --no-configs) holds two more records of
the same run: "dam_break_spheres_128", plain FP32 MGPCG at 128^3 on the
dam-break block union 32 fluid spheres (radius 2-8 cells, seed 1) with its
per-kernel HBM GB/s and a bit-for-bit check of the cell flags and coarse
masks, and "droplets_box_512", the second workload below at 512^3 (two timed
steps per solver).

Second workload (--workload droplets_box, default 512^3): the dam-break block
plus 200 droplets of radius 2-12 cells (seed 3) and a solid 0.1^3 box at a
random pose moving with N(0, 0.3) linear and angular velocity; the face
fractions of the box weight the right-hand side (divergence_fractional_3d).
One step = geometry (face fractions, rigid face velocities), fractional
divergence, solve, gradient subtraction, for the plain FP32 MGPCG, the BF16
mixed solver and the combined learned + BF16 solver; the line reports the
MGPCG fallback of the learned solver, times, HBM GB/s of the geometry and
divergence kernels and the MFMA utilisation of the U-Net.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0
# HBM bytes per launch of the level-0 half-sweeps at 256^3, from two
# rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of `bench.py --kernels-only
# [--kernel ...]`, calibrated on a 67.4 MB device copy
# (profiles/rb_half_sweep_l0_split_traffic_256.md, profiles/rb_half_sweep_l0_traffic_256.md,
# profiles/rb_half_sweep_l0_split_bf16r_traffic_256.md, profiles/p_step_Ap_split_traffic_256.md)
MEASURED_TRAFFIC_L0 = {("rb_half_sweep_l0_split", 256): 153.8e6, ("rb_half_sweep_l0", 256): 242.7e6,
                       ("rb_half_sweep_l0_split_bf16r", 256): 137.4e6, ("p_step_Ap_dot_split", 256): 330.3e6}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["flow_past_sphere", "droplets_box"], default="flow_past_sphere")
    ap.add_argument("--n", type=int, default=0, help="grid size (default 256, or 512 for droplets_box)")
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--solvers", nargs="+", default=["fp32", "fp16", "bf16", "fp32-tuned", "bf16-tuned", "learned"])
    ap.add_argument("--learned-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-judge", action="store_true", help="skip the FP64 CPU oracle projection (error report)")
    ap.add_argument("--no-heldout", action="store_true", help="skip the held-out learned evaluation")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 128^3 dam-break + spheres and the 512^3 droplets + box records")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--kernels-only", action="store_true",
                    help="only replay the level-0 half-sweep (for counter collection)")
    ap.add_argument("--kernel", choices=["split", "split-bf16", "reference-layout", "pstep", "unet", "advect"],
                    default="split",
                    help="what --kernels-only replays: a level-0 half-sweep (FP32 solve, BF16 solve's "
                         "post-smoothing, reference layout), the split solve's one-pass p step with Ap, "
                         "one U-Net forward, or advect_3d")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported"}))
        sys.exit(2)

    from flub_amd import _lib, device, scenes
    from flub_amd.project import Projection, tuned_levels

    _lib.lib()
    if args.workload == "droplets_box":
        return droplets_box_main(args)
    n, L = args.n or 256, args.levels
    t0 = time.time()
    s = scenes.flow_past_sphere(n, 2)
    log(f"scene {n}^3 built in {time.time() - t0:.1f}s; fluid cells {(s['tag'] == 1).sum()}")
    src = [device.from_numpy(s[c]) for c in ("u", "v", "w")]

    if args.kernels_only and args.kernel in ("unet", "advect"):
        return kernels_only_other(args, s, n)
    if args.kernels_only:
        mixed = args.kernel == "split-bf16"
        pr = Projection((n, n, n), n_mg_levels=L, pre_and_post_smoothing=2, bottom_smoothing=6,
                        precision="bf16" if mixed else "fp32")
        pr.load(s)
        pr.mgpcg.set_tags(pr.tag.f)
        # a copy of known size (calibrates FETCH_SIZE / WRITE_SIZE)
        device.copy(pr.vel0.u.f, src[0])
        device.synchronize()
        avg, byt = ctypes.c_float(), ctypes.c_double()
        kid, kname = {"split": (8, "rb_half_sweep_l0_split"), "split-bf16": (12, "rb_half_sweep_l0_split_bf16r"),
                      "reference-layout": (0, "rb_half_sweep_l0"), "pstep": (10, "p_step_Ap_dot_split")}[args.kernel]
        _lib.call("flub_mgpcg_profile", pr.mgpcg.handle, kid, 20, ctypes.byref(avg), ctypes.byref(byt))
        print(json.dumps({"kernel": kname, "ms": avg.value, "bytes": byt.value, "copy_bytes": src[0].nbytes}))
        return

    solvers = {}
    kernels = {}
    mixed_sweep = {}   # per BF16 leg: its level-0 post-smoothing half-sweep replayed
    pstep = {}         # per leg: its one-pass p step with Ap (k_sp_pAp) replayed
    outputs = {}      # per solver: pressure and velocities, for the FP64 judge
    in_solve = None
    last_pr = None
    for name in args.solvers:
        if name == "learned":
            continue
        pr = make_projection(name, n, L)
        pr.load(s)
        dst = [pr.vel0.u.f, pr.vel0.v.f, pr.vel0.w.f]
        times, iters = [], []
        for i in range(args.warmup + args.steps):
            for a, b in zip(dst, src):
                device.copy(a, b)
            device.synchronize()
            t = time.perf_counter()
            pr.project(max_iters=-1, rel_tol=1e-12)
            device.synchronize()
            if i >= args.warmup:
                times.append((time.perf_counter() - t) * 1e3)
                iters.append(pr.iterations)
                log(f"{name} step {i - args.warmup}: {times[-1]:.3f} ms, {pr.iterations} iterations, "
                    f"rTr {pr.rTr:.3e}")
        solvers[name] = {"ms": round(float(np.mean(times)), 4), "ms_min": round(float(np.min(times)), 4),
                         "iterations": int(np.median(iters)), **accuracy(pr, s["tag"], pr.rTr)}
        outputs[name] = (pr.p0.f.numpy(), *pr.velocity(), solution_f64(pr.mgpcg.handle, n))
        if name == "fp32":
            kernels = profile_kernels(pr)
            in_solve = probe_in_solve(pr, src)
        if name.startswith("bf16"):
            mixed_sweep[name] = profile_one(pr, 12)
        pstep[name] = profile_one(pr, 10)
        last_pr = pr

    if "learned" in args.solvers:
        solvers["learned"], outputs["learned"] = learned_leg(s, n, L, src, args.learned_steps, args.warmup)

    if not args.no_judge:
        fp64_judge(s, n, L, outputs, solvers, args.cpu_threads)

    best = min(solvers, key=lambda k: solvers[k]["ms"])
    ms = solvers[best]["ms"]
    # the dominant kernel of the fastest leg (the one-pass p step with Ap, 8
    # launches of some 50 us in a tuned BF16 projection), replayed with HIP
    # events on the solver's stream (it agrees with the rocprofv3 kernel-trace
    # average of those replays); beside it the level-0 half-sweeps: bf16_sweep,
    # the BF16 solve's FP32 post-smoothing sweep with BF16 r (16 launches of
    # some 23 us), and fp32_sweep, the FP32 solve's half-sweep, with in_solve =
    # HIP events around its launches inside a projection's V-cycles (each
    # interval also holds the dispatch after the previous kernel, so it bounds
    # the in-solve kernel time from above).
    dom_name = "rb_half_sweep_l0_split" if "rb_half_sweep_l0_split" in kernels else "rb_half_sweep_l0"
    dom = kernels.get(dom_name)
    fp32_sweep = None
    if dom:
        fp32_sweep = {"kernel": ("k_ssweep level-0 red-black half-sweep, colour-split layout"
                                 if dom_name.endswith("split") else "k_sweep<float> level-0 red-black half-sweep")
                                + " (9 B/cell algorithmic)", "bytes_per_launch": dom["bytes"], "avg_ms": dom["ms"],
                      "achieved": dom["GBs"], "frac": round(dom["GBs"] / HBM_PEAK_GBS, 4),
                      "traffic": MEASURED_TRAFFIC_L0.get((dom_name, n))}
        if in_solve:
            gbs = dom["bytes"] / (in_solve["avg_ms"] * 1e-3) / 1e9
            fp32_sweep["in_solve"] = {**in_solve, "achieved": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    roofline = None
    bf16_sweep = None
    if best in mixed_sweep and mixed_sweep[best]:
        d = mixed_sweep[best]
        bf16_sweep = {"kernel": "k_ssweep<PLAIN, STORE, float, float, bf16> level-0 post-smoothing red-black "
                                "half-sweep of the BF16 solve, colour-split layout (8 B/cell algorithmic: FP32 z, "
                                "BF16 r)", "bytes_per_launch": d["bytes"], "avg_ms": d["ms"], "achieved": d["GBs"],
                      "frac": round(d["GBs"] / HBM_PEAK_GBS, 4),
                      "traffic": MEASURED_TRAFFIC_L0.get(("rb_half_sweep_l0_split_bf16r", n))}
    if pstep.get(best):
        # the kernel with the largest share of the fastest leg's solve: the
        # one-pass p step with Ap (once per PCG iteration, 18 B/cell)
        d = pstep[best]
        roofline = {"bound": "hbm", "achieved": d["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(d["GBs"] / HBM_PEAK_GBS, 4),
                    "traffic": MEASURED_TRAFFIC_L0.get(("p_step_Ap_dot_split", n)),
                    "kernel": "k_sp_pAp: the split solve's p step (p = z + beta p into the ring) with Ap = ndt(p) "
                              "and the FP64 p.Ap partials in one pass, colour-split layout (18 B/cell algorithmic: "
                              "p_old, z, p, Ap FP32, codes)", "leg": best, "bytes_per_launch": d["bytes"],
                    "avg_ms": d["ms"], "bf16_sweep": bf16_sweep, "fp32_sweep": fp32_sweep}
    elif fp32_sweep:
        roofline = {"bound": "hbm", "achieved": fp32_sweep["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": fp32_sweep["frac"], "traffic": fp32_sweep["traffic"], "kernel": fp32_sweep["kernel"],
                    "leg": "fp32", "bytes_per_launch": fp32_sweep["bytes_per_launch"], "avg_ms": fp32_sweep["avg_ms"],
                    "in_solve": fp32_sweep.get("in_solve")}
    heldout = None
    if "learned" in args.solvers and not args.no_heldout:
        heldout = learned_heldout()
    cpu = None if args.no_cpu else cpu_baseline(s, n, L, args.cpu_threads)
    configs = None
    if not args.no_configs:
        del last_pr, outputs
        configs = {"dam_break_spheres_128": dam_break_spheres_128(L, 5, 2),
                   "droplets_box_512": droplets_box(512, L, 2, 1)}
    out = {
        "metric": "ms per pressure projection to 1e-6 rel. residual @256^3: learned/mixed vs MGPCG",
        "value": ms, "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": ("f32" if best.startswith("fp32") else
                  ("bf16+f32" if best.startswith("bf16") or best == "learned" else "f16+f32")),
        "data": "synthetic (seeded scene)",
        "config": {"workload": f"flow past sphere {n}^3: divergence, pressure solve to 1e-6 rel. residual, "
                               f"gradient subtraction; reference MGPCG {L} levels, 2 RB-GS pre/post (doubling per "
                               f"level), 6 bottom; tuned cycle {tuned_levels((n, n, n))} levels, coarse correction "
                               f"x1/2, at most 3 RB-GS per level", "grid": n, "best_solver": best},
        "solvers": solvers,
        "accuracy_rule": "every leg: x in FP64, stop when the FP64 true residual ||b - A x|| < 1e-6 ||b|| "
                         "(the reference's test; r replaced by the true residual and the solve continued "
                         "otherwise)",
        "learned_heldout": heldout,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "configs": configs,
    }
    if "fp32" in solvers:
        out["speedup_vs_mgpcg_fp32"] = {k: round(solvers["fp32"]["ms"] / v["ms"], 3) for k, v in solvers.items()}
    print(json.dumps(out), flush=True)


def make_projection(name, n, L):
    """The projection of a solver leg: "fp32" / "fp16" / "bf16" run the
    reference V-cycle (L levels, 2 pre/post sweeps doubling per level, 6
    bottom); "<precision>-tuned" the tuned cycle (flub_amd.project.TUNED_CYCLE,
    levels down to 4^3)."""
    from flub_amd.project import TUNED_CYCLE, Projection, tuned_levels
    prec, _, variant = name.partition("-")
    if variant == "tuned":
        return Projection((n, n, n), n_mg_levels=tuned_levels((n, n, n)), pre_and_post_smoothing=2,
                          bottom_smoothing=6, precision=prec, **TUNED_CYCLE)
    return Projection((n, n, n), n_mg_levels=L, pre_and_post_smoothing=2, bottom_smoothing=6, precision=prec)


def kernels_only_other(args, s, n, reps=3):
    """For counter collection: `reps` U-Net forwards on an n^3 input (the
    committed weights), or `reps` advect_3d calls (MacCormack advection of the
    three velocity components and a scalar) on the scene's velocity field."""
    import torch
    from flub_amd import device
    if args.kernel == "unet":
        from flub_amd.learned.evaluate import load_net
        from flub_amd.learned.unet import flops_per_voxel
        net = load_net(os.path.join(ROOT, "flub_amd", "learned", "weights", "unet_128.npz"))
        x = torch.randn((1, n, n, n, 2), device="cuda")
        ms = []
        for _ in range(reps):
            t, _ = _timed(lambda: net.infer(x))
            ms.append(t)
        print(json.dumps({"kernel": "unet_forward", "n": n, "ms": ms,
                          "tflops": flops_per_voxel() * n ** 3 / (min(ms) * 1e-3) / 1e12}))
        return
    from flub_amd.advect import advect_3d
    from flub_amd.const_param import Field, make_mac_grid
    vel = make_mac_grid(float, (n, n, n))
    for c, f in zip(("u", "v", "w"), (vel.u, vel.v, vel.w)):
        f.f.assign(s[c])
    phi = Field(f=device.from_numpy(s["phi"]))
    new_vel = make_mac_grid(float, (n, n, n))
    new_q = Field(f=device.empty((n, n, n), float))
    ms = []
    for _ in range(reps):
        t, _ = _timed(lambda: advect_3d(vel, phi, new_vel, new_q, dt=1.0 / 60.0, h0=1.0 / n))
        ms.append(t)
    print(json.dumps({"kernel": "advect_3d", "n": n, "ms": ms, "bytes": 4 * (8 + 12) * n ** 3}))


def _rhs_norm2(pr, tag):
    div = pr.div.f.numpy()
    b = np.where(tag == 1, div * np.float32(-pr.h), np.float32(0)).astype(np.float32)
    return b, float(np.dot(b.astype(np.float64).ravel(), b.astype(np.float64).ravel()))


def true_rel_residual(pr, tag):
    """||b - ndt(x)|| / ||b|| of the FP32 x the last solve handed back,
    evaluated in FP64 on the GPU; b = -h * div on FLUID cells, the right-hand
    side the solve started from."""
    from flub_amd import _lib, device
    b, bb = _rhs_norm2(pr, tag)
    bd = device.from_numpy(b)
    rr = ctypes.c_double()
    _lib.call("flub_mgpcg_residual_norm", pr.mgpcg.handle, pr.mgpcg.x.ptr, bd.ptr, ctypes.byref(rr))
    return float(np.sqrt(rr.value / bb)) if bb > 0 else 0.0


def accuracy(pr, tag, true_rr):
    """The solve's FP64 true relative residual (of its FP64 x, reported by the
    solve) and that of the FP32 x it handed back."""
    _, bb = _rhs_norm2(pr, tag)
    return {"true_rel_res": float(np.sqrt(true_rr / bb)) if bb > 0 else 0.0,
            "true_rel_res_fp32_x": true_rel_residual(pr, tag)}


def fp64_judge(s, n, L, outputs, solvers, threads):
    """The converged FP64 CPU oracle projection of the same scene (rel_tol
    1e-24): each solver's relative pressure error against it (2-norm of the
    FP32 pressure, and the energy norm ||x - x*||_A / ||x*||_A of its FP64 x,
    the error a residual test bounds), and the post-projection divergence of
    each solver's velocity field relative to the divergence before the
    projection (FLUID cells, FP64)."""
    from oracle import oracle as O
    cores = _cpu_cores(threads)
    O.set_threads(cores)
    t = time.perf_counter()
    o = O.MGPCG((n, n, n), L, real=np.float64, pre_and_post_smoothing=2, bottom_smoothing=6)
    _, _, _, p64, _, it64, _ = o.project(s["u"], s["v"], s["w"], s["ur"], s["vr"], s["wr"], s["tag"], 1.0 / n,
                                         rel_tol=1e-24)
    del o
    log(f"FP64 oracle projection: {it64} iterations, {time.perf_counter() - t:.1f} s on {cores} threads")
    fl = s["tag"] == 1
    pre = O.divergence_tag(s["u"], s["v"], s["w"], s["ur"], s["vr"], s["wr"], s["tag"], real=np.float64)
    pre_n = float(np.linalg.norm(pre[fl]))
    pn = float(np.linalg.norm(p64))
    xs = p64 * n   # x* (the pressure is x h, h = 1/n)
    xa = float(np.sum(xs * -O.compute_Ap(xs, s["tag"], real=np.float64)))
    for name, (p, u, v, w, x64) in outputs.items():
        post = O.divergence_tag(u, v, w, s["ur"], s["vr"], s["wr"], s["tag"], real=np.float64)
        solvers[name]["p_rel_err_fp64"] = float(np.linalg.norm(p.astype(np.float64) - p64) / pn)
        e = x64 - xs
        solvers[name]["x_err_energy_fp64"] = float(np.sqrt(abs(np.sum(e * -O.compute_Ap(e, s["tag"], real=np.float64))) / xa))
        solvers[name]["post_div_rel"] = float(np.linalg.norm(post[fl]) / pre_n)
        log(f"{name}: pressure error vs FP64 {solvers[name]['p_rel_err_fp64']:.3e} "
            f"(energy norm {solvers[name]['x_err_energy_fp64']:.3e}), "
            f"post-projection divergence {solvers[name]['post_div_rel']:.3e}")


def probe_in_solve(pr, src):
    """Mean duration of the level-0 plain half-sweep inside one FP32
    projection (HIP events around its launches in the V-cycles)."""
    from flub_amd import _lib, device
    for a, b in zip([pr.vel0.u.f, pr.vel0.v.f, pr.vel0.w.f], src):
        device.copy(a, b)
    device.synchronize()
    try:
        _lib.call("flub_mgpcg_probe", pr.mgpcg.handle, 8, 512)
    except _lib.FlubError as e:
        log(f"in-solve probe: {e}")
        return None
    pr.project(max_iters=-1, rel_tol=1e-12)
    avg, mn, cnt = ctypes.c_float(), ctypes.c_float(), ctypes.c_int()
    _lib.call("flub_mgpcg_probe_read", pr.mgpcg.handle, ctypes.byref(avg), ctypes.byref(mn), ctypes.byref(cnt))
    out = {"avg_ms": round(avg.value, 5), "min_ms": round(mn.value, 5), "launches": cnt.value}
    log(f"in-solve level-0 half-sweep: {out}")
    return out


def profile_one(pr, kid, reps=20):
    """One kernel of a handle replayed with HIP events (flub_mgpcg_profile):
    ms per launch, algorithmic bytes, GB/s."""
    from flub_amd import _lib
    avg, byt = ctypes.c_float(), ctypes.c_double()
    try:
        _lib.call("flub_mgpcg_profile", pr.mgpcg.handle, kid, reps, ctypes.byref(avg), ctypes.byref(byt))
    except _lib.FlubError as e:
        log(f"kernel {kid}: {e}")
        return None
    ent = {"ms": round(avg.value, 5), "bytes": byt.value, "GBs": round(byt.value / (avg.value * 1e-3) / 1e9, 1)}
    log(f"kernel {kid}: {ent}")
    return ent


def profile_kernels(pr):
    """Mean duration of single kernels of the FP32 solve, replayed with HIP
    events on the solver's stream, and their algorithmic bytes."""
    from flub_amd import _lib
    names = {8: "rb_half_sweep_l0_split", 9: "restrict_l0_split", 10: "p_step_Ap_dot_split",
             11: "update_r_dot",
             0: "rb_half_sweep_l0", 7: "blocked4_l0", 5: "prolong_sweep_l0", 4: "restrict_l0",
             1: "p_update_Ap_dot", 2: "update_rx_dot", 6: "update_p", 3: "vcycle"}
    kernels = {}
    for kid, name in names.items():
        avg = ctypes.c_float()
        byt = ctypes.c_double()
        try:
            _lib.call("flub_mgpcg_profile", pr.mgpcg.handle, kid, 20, ctypes.byref(avg), ctypes.byref(byt))
        except _lib.FlubError as e:   # e.g. no blocked smoothing for this grid
            log(f"kernel {name}: {e}")
            continue
        ent = {"ms": round(avg.value, 5), "bytes": byt.value}
        if byt.value > 0:
            ent["GBs"] = round(byt.value / (avg.value * 1e-3) / 1e9, 1)
        kernels[name] = ent
        log(f"kernel {name}: {ent}")
    return kernels


def learned_leg(s, n, L, src, steps, warmup):
    import torch
    from flub_amd import device
    from flub_amd.learned.evaluate import load_net
    from flub_amd.learned.fpcg import LearnedPCG
    from flub_amd.learned.unet import flops_per_voxel
    from flub_amd.project import Projection

    path = os.path.join(ROOT, "flub_amd", "learned", "weights", "unet_128.npz")
    net = load_net(path)
    pr = Projection((n, n, n), n_mg_levels=L, pre_and_post_smoothing=2, bottom_smoothing=6)
    pr.load(s)
    fluid = torch.from_numpy(s["tag"] == 1).cuda()
    solver = LearnedPCG(pr.mgpcg, net)
    dst = [pr.vel0.u.f, pr.vel0.v.f, pr.vel0.w.f]
    times, st = [], None
    for i in range(max(1, warmup) + steps):
        for a, b in zip(dst, src):
            device.copy(a, b)
        device.synchronize()
        t = time.perf_counter()
        st = pr.project_learned(solver, fluid)
        device.synchronize()
        if i >= max(1, warmup):
            times.append((time.perf_counter() - t) * 1e3)
            log(f"learned step: {times[-1]:.2f} ms, {st['iterations']} iterations "
                f"({st['net_iters']} net, {st['vcycle_iters']} V-cycle, fallback at {st['fallback_at']})")
    # one network application, timed alone
    r = torch.randn((1, n, n, n, 2), device="cuda")
    net.infer(r)
    torch.cuda.synchronize()
    t = time.perf_counter()
    net.infer(r)
    torch.cuda.synchronize()
    net_ms = (time.perf_counter() - t) * 1e3
    tflops = flops_per_voxel() * n ** 3 / (net_ms * 1e-3) / 1e12
    ent = {"ms": round(float(np.mean(times)), 3), "iterations": st["iterations"], "net_iters": st["net_iters"],
           "vcycle_iters": st["vcycle_iters"], "fallback_at": st["fallback_at"],
           "true_rel_res": float(np.sqrt(st["true_rTr"] / st["rTr0"])),
           "true_rel_res_fp32_x": float(np.sqrt(solver.fp32_true_residual() / st["rTr0"])),
           "unet_forward_ms": round(net_ms, 3), "unet_tflops": round(tflops, 2),
           "unet_mfma_frac": round(tflops / MFMA_BF16_PEAK_TFLOPS, 4)}
    return ent, (pr.p0.f.numpy(), *pr.velocity(), solver.last_x64.cpu().numpy())


def solution_f64(handle, n):
    """The FP64 x of the handle's last solve (flub_mgpcg_solution_f64)."""
    import torch
    from flub_amd import _lib
    x = torch.empty((n, n, n), dtype=torch.float64, device="cuda")
    _lib.call("flub_mgpcg_solution_f64", handle, ctypes.c_void_p(x.data_ptr()))
    torch.cuda.synchronize()
    return x.cpu().numpy()


def learned_heldout(sizes=(128, 256), seeds=(3000,)):
    """The committed network on held-out seeds (>= 3000; trained on 1000-2999):
    iterations, solve ms and FP64 true residual of plain MGPCG, the learned
    flexible PCG (MGPCG fallback) and the same flexible PCG with the V-cycle
    only, all to the same FP64 true residual (flub_amd/learned/evaluate.py)."""
    from flub_amd.learned.evaluate import evaluate, load_net
    net = load_net(os.path.join(ROOT, "flub_amd", "learned", "weights", "unet_128.npz"))
    out = {}
    for n in sizes:
        res = evaluate(net, n, list(seeds), log=log)
        out[str(n)] = res
    return out


def _timed(fn):
    from flub_amd import device
    device.synchronize()
    t = time.perf_counter()
    out = fn()
    device.synchronize()
    return (time.perf_counter() - t) * 1e3, out


def droplets_box_main(args):
    print(json.dumps(droplets_box(args.n or 512, args.levels, max(1, args.steps), args.warmup)), flush=True)


def droplets_box(n, L, steps, warmup):
    """The dam-break + droplets + moving box workload at n^3 (module
    docstring): one record with every solver's time, the geometry,
    divergence, extrapolation and advection kernels and the U-Net."""
    import torch
    from flub_amd import device, scenes
    from flub_amd.learned.evaluate import load_net
    from flub_amd.learned.fpcg import LearnedPCG
    from flub_amd.learned.unet import flops_per_voxel
    from flub_amd.project import Projection

    t0 = time.time()
    s = scenes.dam_break_droplets_box(n, 3)
    tag = s["tag"]
    log(f"scene {n}^3 built in {time.time() - t0:.1f}s; fluid {(tag == 1).mean():.3f}, solid cells {(tag == 2).sum()}")
    src = [device.from_numpy(s[c]) for c in ("u", "v", "w")]
    fluid = torch.from_numpy(tag == 1).cuda()
    net = load_net(os.path.join(ROOT, "flub_amd", "learned", "weights", "unet_128.npz"))
    out = {}
    legs = ("fp32", "bf16", "bf16-tuned", "learned+bf16")
    for name in legs:
        learned = name.startswith("learned")
        pr = make_projection(name.split("+")[-1], n, L)
        pr.load(s)
        pr.set_bodies(s["bodies"])
        solver = LearnedPCG(pr.mgpcg, net) if learned else None
        dst = [pr.vel0.u.f, pr.vel0.v.f, pr.vel0.w.f]
        times, st = [], None
        for i in range(max(1, warmup) + steps):
            for a, b in zip(dst, src):
                device.copy(a, b)
            ms, st = _timed(lambda: pr.project_fractional(solver, fluid if learned else None, rel_tol=1e-12))
            if i >= max(1, warmup):
                times.append(ms)
                log(f"{name} step: {ms:.2f} ms, " + (f"{st['iterations']} iterations ({st['net_iters']} net, "
                    f"{st['vcycle_iters']} MGPCG, fallback at {st['fallback_at']})" if learned
                    else f"{st} iterations"))
        ent = {"ms": round(float(np.mean(times)), 3), "iterations": st["iterations"] if learned else st,
               **accuracy(pr, tag, pr.rTr)}
        if learned:
            ent.update(net_iters=st["net_iters"], mgpcg_iters=st["vcycle_iters"], fallbacks=int(st["fallback_at"] is not None),
                       fallback_at=st["fallback_at"], handovers=st["handovers"], replacements=st["replacements"])
        out[name] = ent
        if name == "fp32":
            # the geometry and divergence kernels alone
            g_ms, _ = _timed(pr.precompute_geometry)
            d_ms, _ = _timed(pr.divergence_fractional)
            faces = 3 * (n + 1) * n * n
            # fractions + owners written (8 B/face), solid velocities read owner + write (8 B/face)
            out["geometry"] = {"ms": round(g_ms, 3), "GBs": round(16 * faces / (g_ms * 1e-3) / 1e9, 1),
                               "bytes": 16 * faces}
            # fractional divergence: 9 face arrays read + div written (40 B/cell); mask_with_tag
            # (div and int32 tag read, div written: 12 B/cell); scale_by_3d (8 B/cell)
            dbytes = (40 + 12 + 8) * n ** 3
            out["divergence_fractional"] = {"ms": round(d_ms, 3), "GBs": round(dbytes / (d_ms * 1e-3) / 1e9, 1),
                                            "bytes": dbytes}
            out["kernels"] = profile_kernels(pr)
            # the step after the projection: velocity extrapolation into the air
            # (set_unknown + 16 n / 128 passes, computed as one launch per face array:
            # later passes repeat the second, which only adds the faces the first flushed)
            from flub_amd.const_param import Field
            from flub_amd.extrapolate import extrapolate
            phi = Field(f=device.from_numpy(s["phi"]))
            extrapolate(pr.vel0, pr.vel_owner, phi)
            e_ms, _ = _timed(lambda: extrapolate(pr.vel0, pr.vel_owner, phi))
            # set_unknown: phi read + 3 marks written (16 B/cell); the fill: the mark
            # (4 B) of every face, the velocity written on faces to fill, the known
            # neighbours' values read (4 + 4 B/face at most)
            ebytes = 16 * n ** 3 + 12 * faces
            out["extrapolate"] = {"ms": round(e_ms, 3), "GBs": round(ebytes / (e_ms * 1e-3) / 1e9, 1),
                                  "bytes": ebytes, "passes": 16 * n // 128}
            # the step before the projection: MacCormack advection of the velocity
            # and a scalar (advect_3d), out of place
            from flub_amd.advect import advect_3d
            from flub_amd.const_param import make_mac_grid
            new_vel = make_mac_grid(float, (n, n, n))
            new_q = Field(f=device.empty((n, n, n), float))
            advect_3d(pr.vel0, phi, new_vel, new_q, dt=1.0 / 60.0, h0=1.0 / n)
            a_ms, _ = _timed(lambda: advect_3d(pr.vel0, phi, new_vel, new_q, dt=1.0 / 60.0, h0=1.0 / n))
            # compulsory bytes: each of the 4 fields read and written once, the
            # velocity read by the traces (12 B/cell) for each of the 4
            abytes = 4 * (8 + 12) * n ** 3
            out["advect"] = {"ms": round(a_ms, 3), "GBs": round(abytes / (a_ms * 1e-3) / 1e9, 1),
                             "bytes": abytes, "fields": 4}
            del new_vel, new_q
        del pr, solver
    # one network application at this size
    r = torch.randn((1, n, n, n, 2), device="cuda")
    net.infer(r)
    net_ms, _ = _timed(lambda: net.infer(r))
    tflops = flops_per_voxel() * n ** 3 / (net_ms * 1e-3) / 1e12
    out["unet"] = {"forward_ms": round(net_ms, 2), "tflops": round(tflops, 1),
                   "mfma_frac": round(tflops / MFMA_BF16_PEAK_TFLOPS, 4)}
    del r
    torch.cuda.empty_cache()
    best = min(legs, key=lambda k: out[k]["ms"])
    return {"metric": f"ms per pressure projection to 1e-6 rel. residual @{n}^3 with a moving solid box "
                      f"(face-fraction RHS): learned+mixed vs MGPCG",
            "value": out[best]["ms"], "unit": "ms", "n_gpus": 1, "steps": steps, "warmup": warmup,
            "higher_is_better": False, "data": "synthetic (seeded scene)",
            "config": {"workload": f"dam break + 200 droplets + rotated moving box {n}^3: face fractions, "
                                   f"rigid face velocities, fractional divergence, solve, gradient "
                                   f"subtraction", "grid": n, "best_solver": best},
            "solvers": {k: out[k] for k in legs},
            "speedup_vs_mgpcg_fp32": {k: round(out["fp32"]["ms"] / out[k]["ms"], 3) for k in legs},
            "geometry": out["geometry"], "divergence_fractional": out["divergence_fractional"],
            "extrapolate": out["extrapolate"], "advect": out["advect"],
            "unet": out["unet"], "kernels": out["kernels"]}


def dam_break_spheres_128(L, steps, warmup):
    """Plain FP32 MGPCG at 128^3 on the dam-break block union 32 fluid spheres
    (radius 2-8 cells, seed 1), fluid faces N(0, 0.5): ms per projection,
    iterations, FP64 true residual, the per-kernel HBM GB/s of the solve's
    kernels at this size, and whether the GPU's cell flags and coarse masks
    (tag_dirichlet / tag_neumann, set_tags) equal the oracle's bit for bit
    (the oracle only as the checker)."""
    from flub_amd import device, scenes
    from flub_amd.const_param import Field, IntField
    from flub_amd.geometry import tag_dirichlet, tag_neumann
    from flub_amd.project import Projection
    from oracle import oracle as O
    n = 128
    s = scenes.dam_break_spheres(n, 1)
    src = [device.from_numpy(s[c]) for c in ("u", "v", "w")]
    pr = Projection((n, n, n), n_mg_levels=L, pre_and_post_smoothing=2, bottom_smoothing=6)
    pr.load(s)
    dst = [pr.vel0.u.f, pr.vel0.v.f, pr.vel0.w.f]
    times, iters = [], []
    for i in range(warmup + steps):
        for a, b in zip(dst, src):
            device.copy(a, b)
        ms, it = _timed(lambda: pr.project(max_iters=-1, rel_tol=1e-12))
        if i >= warmup:
            times.append(ms)
            iters.append(it)
    ent = {"ms": round(float(np.mean(times)), 4), "ms_min": round(float(np.min(times)), 4),
           "iterations": int(np.median(iters)), **accuracy(pr, s["tag"], pr.rTr)}
    ent["kernels"] = profile_kernels(pr)
    # flags from the level sets on the GPU, and the coarse masks of set_tags
    tag = IntField(f=device.zeros((n, n, n), int))
    tag_dirichlet(tag, Field(f=device.from_numpy(s["phi"])))
    tag_neumann(tag, Field(f=device.from_numpy(s["phi_rigid"])))
    t = O.tag_neumann(s["phi_rigid"], O.tag_dirichlet(s["phi"]))
    flags_ok = bool(np.array_equal(tag.f.numpy(), t))
    pr.mgpcg.set_tags(tag.f)
    masks_ok = True
    for l in range(L):
        masks_ok &= bool(np.array_equal(pr.mgpcg.tag[l].numpy().astype(np.int32), t))
        masks_ok &= bool(np.array_equal(pr.mgpcg.divd[l].numpy().view(np.uint32),
                                        O.compute_divd(t).view(np.uint32)))
        if l + 1 < L:
            t = O.restrict_tag(t)
    ent["flags_bitexact"] = flags_ok
    ent["coarse_masks_and_divd_bitexact"] = masks_ok
    ent["fluid_cells"] = int((s["tag"] == 1).sum())
    log(f"128^3 dam break + 32 spheres: {ent['ms']} ms, {ent['iterations']} iterations, flags {flags_ok}, "
        f"masks {masks_ok}")
    return ent


def _cpu_cores(threads):
    """The threads the CPU legs use: --cpu-threads, else OMP_NUM_THREADS (the
    host's CPU share where it is set), else every core this process may run on."""
    if threads:
        return int(threads)
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_baseline(s, n, L, threads, reps=5):
    """The FP32 CPU oracle (a C restatement of the reference path: the
    reference's kernels in launch order, FP32, red-black colour order) timed on
    the same scene, two ways: OpenMP on every core of the host share -- the
    median of `reps` full projections after one warm-up -- and on one thread
    (as Warp's serial CPU backend runs a launch), from a bounded sample: the
    median of 3 projections capped at 1 and at 2 iterations, extrapolated to
    the full iteration count K as t1 + (K - 1) (t2 - t1)."""
    from oracle import oracle as O
    cores = _cpu_cores(threads)

    def run(max_iters, parallel):
        o = O.MGPCG((n, n, n), L, pre_and_post_smoothing=2, bottom_smoothing=6, parallel_dot=parallel)
        a = time.perf_counter()
        r = o.project(s["u"], s["v"], s["w"], s["ur"], s["vr"], s["wr"], s["tag"], 1.0 / n, rel_tol=1e-12,
                      max_iters=max_iters)
        return (time.perf_counter() - a) * 1e3, r[5]

    O.set_threads(cores)
    run(-1, True)
    multi = [run(-1, True) for _ in range(reps)]
    ms = float(np.median([m for m, _ in multi]))
    K = multi[0][1]
    log(f"cpu oracle: median {ms:.1f} ms of {reps} projections, {K} iterations, {cores} threads")
    O.set_threads(1)
    t1 = float(np.median([run(1, False)[0] for _ in range(3)]))
    t2 = float(np.median([run(2, False)[0] for _ in range(3)]))
    single = t1 + (K - 1) * (t2 - t1)
    O.set_threads(cores)
    log(f"cpu oracle, 1 thread: {t1:.0f} ms (1 iteration), {t2:.0f} ms (2) -> {single:.0f} ms for {K}")
    return {"value": round(ms, 2), "unit": "ms", "cores": cores, "kind": "port",
            "sample": f"median of {reps} full projections of the same {n}^3 scene (after one warm-up) by the "
                      f"FP32 C oracle, OpenMP on {cores} threads, {K} iterations",
            "single_thread": {"value": round(single, 1), "unit": "ms", "cores": 1,
                              "sample": f"median of 3 projections capped at 1 and at 2 iterations "
                                        f"({t1:.0f} and {t2:.0f} ms), extrapolated to {K} iterations"}}


if __name__ == "__main__":
    main()
