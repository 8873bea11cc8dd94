"""Train the U-Net preconditioner on the GPU with the project's conv kernels.

    python -m flub_amd.learned.train --n 128 --pairs 2000 --steps 1500 --out weights/unet_128.npz

Pairs come from flub_amd.learned.data (seeds 1000..1000+pairs-1, solved by
the library's FP32 MGPCG).  Loss: relative squared error of the predicted
A^-1 F over FLUID cells, Adam.  Prints one line per log interval.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np
import torch

from .data import make_pairs
from .unet import UNet

HERE = os.path.dirname(os.path.abspath(__file__))


def rel_loss(pred, target, fluid):
    num = ((pred - target) ** 2 * fluid).sum(dim=(1, 2, 3))
    den = (target ** 2 * fluid).sum(dim=(1, 2, 3)).clamp_min(1e-30)
    return (num / den).mean()


def train(n=128, pairs=2000, steps=1500, batch=2, lr=1e-3, seed=0, first_seed=1000, log=print, out=None):
    torch.manual_seed(seed)
    dev = "cuda"
    t0 = time.time()
    X, Y = make_pairs(n, list(range(first_seed, first_seed + pairs)), log=log)
    log(f"generated {pairs} pairs at {n}^3 in {time.time() - t0:.1f}s")
    # fixed output scale: the network predicts Y / yscale (O(1) values)
    yscale = float(torch.sqrt((Y.float() ** 2).sum() / X[..., 1].float().sum()))
    log(f"target scale {yscale:.3f}")
    net = UNet(seed).to(dev)
    net.out_scale = yscale
    opt = torch.optim.Adam(net.parameters(), lr=lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, T_max=steps, eta_min=lr * 0.05)
    g = torch.Generator().manual_seed(seed)
    hist = []
    t1 = time.time()
    for step in range(steps):
        idx = torch.randint(0, pairs, (batch,), generator=g)
        xb = X[idx].float()
        yb = Y[idx]
        fl = xb[..., 1]
        pred = net(xb) * fl
        loss = rel_loss(pred, yb / yscale, fl)
        opt.zero_grad(set_to_none=True)
        loss.backward()
        torch.nn.utils.clip_grad_norm_(net.parameters(), 1.0)
        opt.step()
        sched.step()
        hist.append(loss.item())
        if step % 50 == 0 or step == steps - 1:
            log(f"step {step}: loss {np.mean(hist[-50:]):.4f} ({(time.time() - t1) / (step + 1) * 1e3:.1f} ms/step)")
    if out:
        os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
        np.savez(out, **net.state_numpy(), n=np.int32(n), pairs=np.int32(pairs), steps=np.int32(steps),
                 final_loss=np.float32(np.mean(hist[-50:])), loss_history=np.float32(hist))
        log(f"saved {out}")
    return net, hist


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--pairs", type=int, default=2000)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--batch", type=int, default=2)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=os.path.join(HERE, "weights", "unet_128.npz"))
    a = ap.parse_args()
    train(a.n, a.pairs, a.steps, a.batch, a.lr, out=a.out, log=lambda s: print(s, flush=True))


if __name__ == "__main__":
    sys.exit(main())
