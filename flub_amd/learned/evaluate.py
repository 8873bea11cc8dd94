"""Evaluate the learned preconditioner on held-out scenes (seeds >= 3000):
iterations and solve time of (a) the plain FP32 MGPCG of the library, (b) the
learned flexible PCG with V-cycle fallback, (c) the same flexible PCG with the
V-cycle only, (d) MGPCG with the tuned V-cycle (flub_amd.project.TUNED_CYCLE)
and (e) the learned flexible PCG falling back to that tuned MGPCG, all to the
same FP64 true relative residual (1e-6: x in FP64, the true residual
recomputed and replacing r when the recursive one passes).

    python -m flub_amd.learned.evaluate --weights weights/unet_128.npz --n 128 --seeds 3000 3001
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np
import torch

from ..multigrid.mgpcg import MGPCG
from ..project import TUNED_CYCLE, tuned_levels
from .data import _dev, divergence, scene_torch
from .fpcg import LearnedPCG
from .unet import UNet

HERE = os.path.dirname(os.path.abspath(__file__))


def load_net(path):
    net = UNet(0).cuda()
    with np.load(path, allow_pickle=False) as f:
        net.load_numpy({k: f[k] for k in f.files})
    return net


def evaluate(net, n, seeds, levels=4, rel_tol=1e-12, log=print):
    out = []
    mg = MGPCG(3, (n, n, n), levels)
    mg.pre_and_post_smoothing = 2
    mg.bottom_smoothing = 6
    solver = LearnedPCG(mg, net)
    mgt = MGPCG(3, (n, n, n), tuned_levels((n, n, n)))
    mgt.pre_and_post_smoothing = 2
    mgt.bottom_smoothing = 6
    mgt.coarse_scale = TUNED_CYCLE["coarse_scale"]
    mgt.sweep_cap = TUNED_CYCLE["sweep_cap"]
    solver_t = LearnedPCG(mgt, net)
    for sd in seeds:
        tag, u, v, w = scene_torch(n, sd)
        F = divergence(tag, u, v, w)
        fluid = tag == 1
        mg.set_tags(_dev(tag))
        # (a) plain MGPCG
        res = {}
        for rep in range(2):
            torch.cuda.synchronize()
            t = time.perf_counter()
            mg.init(_dev(F), 1.0)
            mg.mask_interior()
            it = mg.solve(-1, abs_tol=0.0, rel_tol=rel_tol, smoother="phases")
            torch.cuda.synchronize()
            res["mgpcg"] = {"iterations": it, "ms": (time.perf_counter() - t) * 1e3}
        rr0 = float(torch.dot(F.view(-1).double(), F.view(-1).double()))
        # the FP64 true residual of the solve's FP64 x (the split solves' stopping rule)
        res["mgpcg"]["true_rel_res"] = float(np.sqrt(mg.last_rTr / rr0)) if rr0 > 0 else 0.0
        # (b) learned with fallback, (c) flexible PCG, V-cycle only
        for name, use_net in (("learned", True), ("fpcg_vcycle", False)):
            for rep in range(2):
                x, st = solver.solve(F, fluid, rel_tol=rel_tol, use_net=use_net)
            res[name] = {k: st[k] for k in ("iterations", "ms", "net_iters", "vcycle_iters", "fallback_at")}
            res[name]["true_rel_res"] = float(np.sqrt(st["true_rTr"] / st["rTr0"]))
        # (d) MGPCG with the tuned cycle, (e) learned falling back to it
        mgt.set_tags(_dev(tag))
        for rep in range(2):
            torch.cuda.synchronize()
            t = time.perf_counter()
            mgt.init(_dev(F), 1.0)
            mgt.mask_interior()
            it = mgt.solve(-1, abs_tol=0.0, rel_tol=rel_tol, smoother="phases")
            torch.cuda.synchronize()
            res["mgpcg_tuned"] = {"iterations": it, "ms": (time.perf_counter() - t) * 1e3}
        res["mgpcg_tuned"]["true_rel_res"] = float(np.sqrt(mgt.last_rTr / rr0)) if rr0 > 0 else 0.0
        for rep in range(2):
            x, st = solver_t.solve(F, fluid, rel_tol=rel_tol, use_net=True)
        res["learned_tuned"] = {k: st[k] for k in ("iterations", "ms", "net_iters", "vcycle_iters", "fallback_at")}
        res["learned_tuned"]["true_rel_res"] = float(np.sqrt(st["true_rTr"] / st["rTr0"]))
        res["seed"] = sd
        out.append(res)
        log(json.dumps(res))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--weights", default=os.path.join(HERE, "weights", "unet_128.npz"))
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--seeds", type=int, nargs="+", default=[3000, 3001, 3002])
    a = ap.parse_args()
    net = load_net(a.weights)
    evaluate(net, a.n, a.seeds, log=lambda s: print(s, flush=True))


if __name__ == "__main__":
    main()
