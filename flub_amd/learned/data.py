"""Training/evaluation pairs for the learned preconditioner, generated on the
GPU: seeded random scenes (water level U[0.2, 0.7], 1-40 FLUID spheres and
0-4 SOLID spheres of radius U[3, 20] cells, fluid-face velocities N(0, 0.5))
are classified, their divergence F taken (flub_divergence_tag + mask) and
solved by the plain FP32 MGPCG of this library to 1e-6 relative residual.
A pair is (F, y) with y = A^-1 F = -x (x the solver's unknown, which solves
neighbor_dot_tag(x) = F), normalised by the RMS of F over FLUID cells.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _lib
from ..device import DeviceArray
from ..multigrid.mgpcg import MGPCG


def _dev(t):
    return DeviceArray(tuple(t.shape), {torch.float32: np.float32, torch.int32: np.int32}[t.dtype],
                       ptr=t.data_ptr(), owner=t)


def scene_torch(n: int, seed: int, device="cuda"):
    """(tag int32 (n,n,n), u, v, w float32 face fields) of one random pool scene."""
    rng = np.random.default_rng(seed)
    level = rng.uniform(0.2, 0.7)
    nf = int(rng.integers(1, 41))
    ns = int(rng.integers(0, 5))
    cf = rng.uniform(0.0, 1.0, (nf, 3))
    rf = rng.uniform(3.0, 20.0, nf) / n
    cs = rng.uniform(0.0, 1.0, (ns, 3))
    rs = rng.uniform(3.0, 20.0, ns) / n
    c = (torch.arange(n, device=device, dtype=torch.float32) + 0.5) / n
    x, y, z = c.view(n, 1, 1), c.view(1, n, 1), c.view(1, 1, n)
    phi = (y - float(level)).expand(n, n, n).clone()
    for (a, b, cc), r in zip(cf, rf):
        phi = torch.minimum(phi, torch.sqrt((x - a) ** 2 + (y - b) ** 2 + (z - cc) ** 2) - float(r))
    tag = torch.where(phi > 0, 0, 1).to(torch.int32)
    for (a, b, cc), r in zip(cs, rs):
        tag[torch.sqrt((x - a) ** 2 + (y - b) ** 2 + (z - cc) ** 2) - float(r) < 0] = 2
    g = torch.Generator(device=device).manual_seed(int(seed))
    u = torch.randn((n + 1, n, n), device=device, generator=g) * 0.5
    v = torch.randn((n, n + 1, n), device=device, generator=g) * 0.5
    w = torch.randn((n, n, n + 1), device=device, generator=g) * 0.5
    fl = tag == 1
    mu = torch.zeros_like(u, dtype=torch.bool)
    mv = torch.zeros_like(v, dtype=torch.bool)
    mw = torch.zeros_like(w, dtype=torch.bool)
    mu[1:-1] = fl[1:] | fl[:-1]
    mv[:, 1:-1] = fl[:, 1:] | fl[:, :-1]
    mw[:, :, 1:-1] = fl[:, :, 1:] | fl[:, :, :-1]
    return tag, u * mu, v * mv, w * mw


def divergence(tag, u, v, w):
    """Flux sum F (divergence_tag + mask_with_tag), zero rigid velocities."""
    n = tag.shape
    F = torch.empty(n, device=tag.device, dtype=torch.float32)
    zu, zv, zw = torch.zeros_like(u), torch.zeros_like(v), torch.zeros_like(w)
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    _lib.call("flub_divergence_tag", p(u), p(v), p(w), p(F), p(zu), p(zv), p(zw), p(tag), *n, None)
    _lib.call("flub_mask_with_tag", p(F), p(tag), F.numel(), None)
    return F


def solve_pair(mg: MGPCG, tag, F, rel_tol=1e-12):
    """Solve neighbor_dot_tag(x) = F with the FP32 MGPCG; returns (x, iterations)."""
    torch.cuda.synchronize()
    mg.set_tags(_dev(tag))
    mg.init(_dev(F), 1.0)
    mg.mask_interior()
    it = mg.solve(-1, abs_tol=0.0, rel_tol=rel_tol, smoother="phases")
    x = torch.empty_like(F)
    _lib.call("flub_memcpy_d2d", x.data_ptr(), mg.x.ptr, mg.x.nbytes, None)
    torch.cuda.synchronize()
    return x, it


def normalise(F, fluid):
    s = torch.sqrt((F[fluid] ** 2).mean()).clamp_min(1e-30)
    return s


def make_pairs(n: int, seeds, levels=4, device="cuda", log=None):
    """Stacked normalised inputs X (S, n, n, n, 2) and targets Y (S, n, n, n)."""
    mg = MGPCG(3, (n, n, n), levels)
    mg.pre_and_post_smoothing = 2
    mg.bottom_smoothing = 6
    X = torch.empty((len(seeds), n, n, n, 2), device=device, dtype=torch.bfloat16)
    Y = torch.empty((len(seeds), n, n, n), device=device, dtype=torch.float32)
    for i, sd in enumerate(seeds):
        tag, u, v, w = scene_torch(n, sd, device)
        F = divergence(tag, u, v, w)
        fl = tag == 1
        x, it = solve_pair(mg, tag, F)
        s = normalise(F, fl)
        X[i, ..., 0] = (F / s).to(torch.bfloat16)
        X[i, ..., 1] = fl.to(torch.bfloat16)
        Y[i] = torch.where(fl, -x / s, torch.zeros_like(x))
        if log and (i % 100 == 0 or i == len(seeds) - 1):
            log(f"pair {i + 1}/{len(seeds)} (seed {sd}): {it} iterations")
    return X, Y
