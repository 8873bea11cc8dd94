"""Benchmark: ms per pressure projection to 1e-6 relative residual on one MI355X.

Workload (default): flow past a sphere on a 256^3 MAC grid -- solid sphere of
radius 0.15 at the centre, inflow u=1 through a SOLID ghost layer at x-min,
outflow through an AIR layer at x-max, u = 1 + N(0, 0.1), v, w = N(0, 0.1)
(seed 2).  One step = one full projection (divergence, MGPCG solve to
||r||/||r0|| < 1e-6, gradient subtraction) with inputs resident in HBM; the
velocity inputs are restored by a device copy outside the timed region
before every step.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--solver fp32]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--solver", default="fp32", choices=["fp32", "fp16", "bf16"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported"}))
        sys.exit(2)

    from flub_amd import _lib, device, scenes
    from flub_amd.project import Projection

    _lib.lib()
    n, L = args.n, args.levels
    t0 = time.time()
    s = scenes.flow_past_sphere(n, 2)
    log(f"scene {n}^3 built in {time.time() - t0:.1f}s; fluid cells {(s['tag'] == 1).sum()}")
    pr = Projection((n, n, n), n_mg_levels=L, pre_and_post_smoothing=2, bottom_smoothing=6,
                    precision=args.solver)
    pr.load(s)
    # pristine copies of the inputs in HBM
    src = [device.from_numpy(s[c]) for c in ("u", "v", "w")]
    dst = [pr.vel0.u.f, pr.vel0.v.f, pr.vel0.w.f]

    def reset():
        for a, b in zip(dst, src):
            device.copy(a, b)

    iters = []
    for i in range(args.warmup):
        reset()
        pr.project(max_iters=-1, rel_tol=1e-12)
        log(f"warmup {i}: {pr.iterations} iterations")
    times = []
    for i in range(args.steps):
        reset()
        device.synchronize()
        a = time.perf_counter()
        pr.project(max_iters=-1, rel_tol=1e-12)
        device.synchronize()
        times.append((time.perf_counter() - a) * 1e3)
        iters.append(pr.iterations)
        log(f"step {i}: {times[-1]:.3f} ms, {pr.iterations} iterations, rTr {pr.rTr:.3e}")
    ms = float(np.mean(times))

    # accuracy of the last projection against the initial residual
    p = pr.p0.f.numpy()

    # per-kernel timing with HIP events on the solver's stream
    kernels = {}
    names = {0: "rb_half_sweep_l0", 1: "Ap_dot", 2: "update_rx_dot", 4: "restrict_l0", 5: "prolong_l0",
             6: "update_p", 3: "vcycle"}
    import ctypes
    for kid, name in names.items():
        avg = ctypes.c_float()
        byt = ctypes.c_double()
        _lib.call("flub_mgpcg_profile", pr.mgpcg.handle, kid, 20, ctypes.byref(avg), ctypes.byref(byt))
        ent = {"ms": round(avg.value, 5), "bytes": byt.value}
        if byt.value > 0:
            ent["GBs"] = round(byt.value / (avg.value * 1e-3) / 1e9, 1)
        kernels[name] = ent
        log(f"kernel {name}: {ent}")
    dom = kernels["rb_half_sweep_l0"]
    roofline = {"bound": "hbm", "achieved": dom["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(dom["GBs"] / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_smooth_mc (level-0 red-black half-sweep)",
                "bytes_per_launch": dom["bytes"]}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(s, n, L, args.cpu_threads)

    out = {
        "metric": "ms per pressure projection to 1e-6 rel. residual @256^3: learned/mixed vs MGPCG",
        "value": round(ms, 4), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.solver == "fp32" else f"{args.solver}+f32",
        "data": "synthetic (seeded scene)",
        "config": {"workload": f"flow past sphere {n}^3, {args.solver} MGPCG ({L} levels, 2 pre/post RBGS, "
                               f"bottom 6), rel. residual 1e-6", "grid": n, "levels": L, "solver": args.solver,
                   "iterations": int(np.median(iters))},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(out), flush=True)


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
