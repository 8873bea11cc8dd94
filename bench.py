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
storage inside FP32 flexible CG with FP64 reductions) and "learned" (U-Net
preconditioner in flexible PCG with V-cycle fallback, weights
flub_amd/learned/weights/unet_128.npz).  `value` is the fastest solver's
ms per projection; every solver's figures are in "solvers".

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--solvers fp32 fp16 bf16 learned]

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
# HBM bytes per launch of the level-0 FP32 half-sweeps at 256^3, from two
# rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of `bench.py --kernels-only
# [--kernel ...]`, calibrated on a 67.4 MB device copy
# (profiles/rb_half_sweep_l0_split_traffic_256.md, profiles/rb_half_sweep_l0_traffic_256.md)
MEASURED_TRAFFIC_L0 = {("rb_half_sweep_l0_split", 256): 153.8e6, ("rb_half_sweep_l0", 256): 242.7e6}


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
    ap.add_argument("--solvers", nargs="+", default=["fp32", "fp16", "bf16", "learned"])
    ap.add_argument("--learned-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--kernels-only", action="store_true",
                    help="only replay the level-0 half-sweep (for counter collection)")
    ap.add_argument("--kernel", choices=["split", "reference-layout"], default="split",
                    help="which level-0 half-sweep --kernels-only replays")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported"}))
        sys.exit(2)

    from flub_amd import _lib, device, scenes
    from flub_amd.project import Projection

    _lib.lib()
    if args.workload == "droplets_box":
        return droplets_box_main(args)
    n, L = args.n or 256, args.levels
    t0 = time.time()
    s = scenes.flow_past_sphere(n, 2)
    log(f"scene {n}^3 built in {time.time() - t0:.1f}s; fluid cells {(s['tag'] == 1).sum()}")
    src = [device.from_numpy(s[c]) for c in ("u", "v", "w")]

    if args.kernels_only:
        pr = Projection((n, n, n), n_mg_levels=L, pre_and_post_smoothing=2, bottom_smoothing=6)
        pr.load(s)
        pr.mgpcg.set_tags(pr.tag.f)
        # a copy of known size (calibrates FETCH_SIZE / WRITE_SIZE)
        device.copy(pr.vel0.u.f, src[0])
        device.synchronize()
        avg, byt = ctypes.c_float(), ctypes.c_double()
        kid, kname = (8, "rb_half_sweep_l0_split") if args.kernel == "split" else (0, "rb_half_sweep_l0")
        _lib.call("flub_mgpcg_profile", pr.mgpcg.handle, kid, 20, ctypes.byref(avg), ctypes.byref(byt))
        print(json.dumps({"kernel": kname, "ms": avg.value, "bytes": byt.value, "copy_bytes": src[0].nbytes}))
        return

    solvers = {}
    kernels = {}
    last_pr = None
    for name in args.solvers:
        if name == "learned":
            continue
        pr = Projection((n, n, n), n_mg_levels=L, pre_and_post_smoothing=2, bottom_smoothing=6, precision=name)
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
                         "iterations": int(np.median(iters)), "true_rel_res": true_rel_residual(pr, s["tag"])}
        if name == "fp32":
            kernels = profile_kernels(pr)
        last_pr = pr

    if "learned" in args.solvers:
        solvers["learned"] = learned_leg(s, n, L, src, args.learned_steps, args.warmup)

    best = min(solvers, key=lambda k: solvers[k]["ms"])
    ms = solvers[best]["ms"]
    # the dominant kernel: the level-0 half-sweep the solve runs (colour-split when available)
    dom_name = "rb_half_sweep_l0_split" if "rb_half_sweep_l0_split" in kernels else "rb_half_sweep_l0"
    dom = kernels.get(dom_name)
    roofline = None
    if dom:
        roofline = {"bound": "hbm", "achieved": dom["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(dom["GBs"] / HBM_PEAK_GBS, 4), "traffic": MEASURED_TRAFFIC_L0.get((dom_name, n)),
                    "kernel": ("k_ssweep level-0 red-black half-sweep, colour-split layout"
                               if dom_name.endswith("split") else "k_sweep<float> level-0 red-black half-sweep")
                              + " (9 B/cell algorithmic)",
                    "bytes_per_launch": dom["bytes"]}
    cpu = None if args.no_cpu else cpu_baseline(s, n, L, args.cpu_threads)
    out = {
        "metric": "ms per pressure projection to 1e-6 rel. residual @256^3: learned/mixed vs MGPCG",
        "value": ms, "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if best in ("fp32",) else ("bf16+f32" if best in ("bf16", "learned") else "f16+f32"),
        "data": "synthetic (seeded scene)",
        "config": {"workload": f"flow past sphere {n}^3: divergence, pressure solve to 1e-6 rel. residual, "
                               f"gradient subtraction; MGPCG {L} levels, 2 RB-GS pre/post (doubling per level), "
                               f"6 bottom", "grid": n, "best_solver": best},
        "solvers": solvers,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    if "fp32" in solvers:
        out["speedup_vs_mgpcg_fp32"] = {k: round(solvers["fp32"]["ms"] / v["ms"], 3) for k, v in solvers.items()}
    print(json.dumps(out), flush=True)


def true_rel_residual(pr, tag):
    """||b - ndt(x)|| / ||b|| of the last solve, evaluated in FP64 on the GPU;
    b = -h * div on FLUID cells, the right-hand side the solve started from."""
    from flub_amd import _lib, device
    div = pr.div.f.numpy()
    b = np.where(tag == 1, div * np.float32(-pr.h), np.float32(0)).astype(np.float32)
    bd = device.from_numpy(b)
    rr = ctypes.c_double()
    _lib.call("flub_mgpcg_residual_norm", pr.mgpcg.handle, pr.mgpcg.x.ptr, bd.ptr, ctypes.byref(rr))
    bb = float(np.dot(b.astype(np.float64).ravel(), b.astype(np.float64).ravel()))
    return float(np.sqrt(rr.value / bb)) if bb > 0 else 0.0


def profile_kernels(pr):
    """Mean duration of single kernels of the FP32 solve, replayed with HIP
    events on the solver's stream, and their algorithmic bytes."""
    from flub_amd import _lib
    names = {8: "rb_half_sweep_l0_split", 9: "restrict_l0_split", 10: "px_update_then_Ap_dot_split",
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
    return {"ms": round(float(np.mean(times)), 3), "iterations": st["iterations"], "net_iters": st["net_iters"],
            "vcycle_iters": st["vcycle_iters"], "fallback_at": st["fallback_at"],
            "true_rel_res": float(np.sqrt(st["true_rTr"] / st["rTr0"])),
            "unet_forward_ms": round(net_ms, 3), "unet_tflops": round(tflops, 2),
            "unet_mfma_frac": round(tflops / MFMA_BF16_PEAK_TFLOPS, 4)}


def _timed(fn):
    from flub_amd import device
    device.synchronize()
    t = time.perf_counter()
    out = fn()
    device.synchronize()
    return (time.perf_counter() - t) * 1e3, out


def droplets_box_main(args):
    import torch
    from flub_amd import device, scenes
    from flub_amd.learned.evaluate import load_net
    from flub_amd.learned.fpcg import LearnedPCG
    from flub_amd.learned.unet import flops_per_voxel
    from flub_amd.project import Projection

    n, L = args.n or 512, args.levels
    t0 = time.time()
    s = scenes.dam_break_droplets_box(n, 3)
    tag = s["tag"]
    log(f"scene {n}^3 built in {time.time() - t0:.1f}s; fluid {(tag == 1).mean():.3f}, solid cells {(tag == 2).sum()}")
    src = [device.from_numpy(s[c]) for c in ("u", "v", "w")]
    fluid = torch.from_numpy(tag == 1).cuda()
    net = load_net(os.path.join(ROOT, "flub_amd", "learned", "weights", "unet_128.npz"))
    out = {}
    steps = max(1, args.steps)
    for name, prec, learned in (("fp32", "fp32", False), ("bf16", "bf16", False), ("learned+bf16", "bf16", True)):
        pr = Projection((n, n, n), n_mg_levels=L, pre_and_post_smoothing=2, bottom_smoothing=6, precision=prec)
        pr.load(s)
        pr.set_bodies(s["bodies"])
        solver = LearnedPCG(pr.mgpcg, net) if learned else None
        dst = [pr.vel0.u.f, pr.vel0.v.f, pr.vel0.w.f]
        times, st = [], None
        for i in range(max(1, args.warmup) + steps):
            for a, b in zip(dst, src):
                device.copy(a, b)
            ms, st = _timed(lambda: pr.project_fractional(solver, fluid if learned else None, rel_tol=1e-12))
            if i >= max(1, args.warmup):
                times.append(ms)
                log(f"{name} step: {ms:.2f} ms, " + (f"{st['iterations']} iterations ({st['net_iters']} net, "
                    f"{st['vcycle_iters']} MGPCG, fallback at {st['fallback_at']})" if learned
                    else f"{st} iterations"))
        ent = {"ms": round(float(np.mean(times)), 3), "iterations": st["iterations"] if learned else st,
               "true_rel_res": true_rel_residual(pr, tag)}
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
    best = min(("fp32", "bf16", "learned+bf16"), key=lambda k: out[k]["ms"])
    print(json.dumps({"metric": f"ms per pressure projection to 1e-6 rel. residual @{n}^3 with a moving solid box "
                                f"(face-fraction RHS): learned+mixed vs MGPCG",
                      "value": out[best]["ms"], "unit": "ms", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
                      "higher_is_better": False, "data": "synthetic (seeded scene)",
                      "config": {"workload": f"dam break + 200 droplets + rotated moving box {n}^3: face fractions, "
                                             f"rigid face velocities, fractional divergence, solve, gradient "
                                             f"subtraction", "grid": n, "best_solver": best},
                      "solvers": {k: out[k] for k in ("fp32", "bf16", "learned+bf16")},
                      "geometry": out["geometry"], "divergence_fractional": out["divergence_fractional"],
                      "extrapolate": out["extrapolate"], "advect": out["advect"],
                      "unet": out["unet"], "kernels": out["kernels"]}), flush=True)


def cpu_baseline(s, n, L, threads):
    """The FP32 CPU oracle (a C restatement of the reference path, OpenMP) timed
    on one full projection of the same scene."""
    from oracle import oracle as O
    cores = threads or min(len(os.sched_getaffinity(0)), 16)
    O.set_threads(cores)
    o = O.MGPCG((n, n, n), L, pre_and_post_smoothing=2, bottom_smoothing=6, parallel_dot=True)
    a = time.perf_counter()
    r = o.project(s["u"], s["v"], s["w"], s["ur"], s["vr"], s["wr"], s["tag"], 1.0 / n, rel_tol=1e-12)
    ms = (time.perf_counter() - a) * 1e3
    log(f"cpu oracle: {ms:.1f} ms, {r[5]} iterations on {cores} threads")
    return {"value": round(ms, 2), "unit": "ms", "cores": cores, "kind": "port",
            "sample": f"one full projection of the same {n}^3 scene by the FP32 C oracle (OpenMP), "
                      f"{r[5]} iterations"}


if __name__ == "__main__":
    main()
