"""Flexible PCG with the learned preconditioner and MGPCG as fallback and judge.

Solves neighbor_dot_tag(x) = b (the reference's system, mgpcg.py:378-462,
same sign conventions: q = neighbor_dot_tag(p) = -A p, the preconditioner
approximates A^-1) with the Polak-Ribiere beta of flexible CG, FP64
reductions and the reference's squared-norm stopping test r.r < max(abs_tol,
rel_tol * r0.r0).

The solve starts on the network, z = s * U-Net([r/s, FLUID]) (s = RMS of r
on FLUID cells), and falls back to MGPCG for good as soon as one iteration
reduces |r| by less than the factor `stagnation` (MGPCG as automatic
fallback).  The fallback hands the current residual to the library's fused
MGPCG solve on the same handle (FP32, or the FP16/BF16 V-cycle of a mixed
handle) with the absolute tolerance of the original test, and adds its FP64
solution to x; with handover=False it continues this loop with the V-cycle
as preconditioner instead.  x accumulates in FP64, as in the library's
colour-split solves.  When the recursive residual passes the test, the FP64
true residual b - A x is recomputed with the library's operator (MGPCG as
judge of the error); the solve stops if it passes the same test, else it
replaces r and the solve continues on MGPCG (the library solves' rule).  The
returned x is the FP64 solution rounded to FP32.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np
import torch

from .. import _lib


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


class LearnedPCG:
    def __init__(self, mg, net, stagnation=0.5, max_net_iters=1 << 30, max_replacements=3):
        self.mg = mg
        self.net = net
        self.stagnation = stagnation
        self.max_net_iters = max_net_iters
        self.max_replacements = max_replacements

    def apply_A(self, x, y):
        _lib.call("flub_mgpcg_apply_A", self.mg.handle, _p(x), _p(y))

    def true_residual(self, x64, b, r_out=None):
        """||b - A x||^2 over the FLUID cells with the FP64 stencil from the FP64
        x; r_out (FP32, optional) receives b - A x."""
        rr = ctypes.c_double()
        _lib.call("flub_mgpcg_true_residual_f64", self.mg.handle, _p(x64), _p(b),
                  _p(r_out) if r_out is not None else None, ctypes.byref(rr))
        return rr.value

    def vcycle(self, r, z):
        _lib.call("flub_mgpcg_precondition_into", self.mg.handle, _p(r), _p(z), 1)

    def net_precondition(self, r, fluid):
        """z = s * U-Net([r/s, FLUID]) with s the RMS of r over the FLUID cells (r
        is 0 elsewhere)."""
        if getattr(self, "_fluid_ref", None) is not fluid:
            self._fluid_ref = fluid
            self._fluid_u8 = fluid.to(torch.uint8).contiguous()
            self._n_fluid = max(int(fluid.sum()), 1)
        s = math.sqrt(max(float(torch.dot(r.view(-1), r.view(-1))), 0.0) / self._n_fluid)
        if not s > 0.0:
            return torch.zeros_like(r)
        return self.net.precondition(r, self._fluid_u8, s)

    def dot(self, a, b):
        """a.b of two contiguous FP32 device tensors, accumulated in FP64 by the
        library's two-stage reduction (flub_inner_noalloc: its partials live in
        a static device array; the solve's calls run one after another on the
        null stream)."""
        out = getattr(self, "_dot_out", None)
        if out is None or out.device != a.device:
            out = self._dot_out = torch.zeros(1, dtype=torch.float64, device=a.device)
        _lib.call("flub_inner_noalloc", _p(a), _p(b), a.numel(), _p(out), None)
        return float(out.item())

    def _as_device(self, t):
        from ..device import DeviceArray
        return DeviceArray(tuple(t.shape), float, ptr=t.data_ptr(), owner=t)

    def mg_solve(self, r, tol, max_iters):
        """The library's fused MGPCG on residual r down to r.r < tol; returns
        (e, iterations) with e its FP64 solution."""
        self.mg.init(self._as_device(r), 1.0)
        self.mg.mask_interior()
        its = self.mg.solve(max_iters, abs_tol=tol, rel_tol=0.0, smoother="phases")
        e = torch.empty(r.shape, dtype=torch.float64, device=r.device)
        _lib.call("flub_mgpcg_solution_f64", self.mg.handle, _p(e))
        return e, its

    def solve(self, b, fluid, max_iters=-1, rel_tol=1e-12, abs_tol=0.0, eps=1e-12, use_net=True, handover=True):
        """b: FP32 right-hand side (FLUID-masked); fluid: bool mask.  Returns
        (x, stats) with x the FP32 solution (the FP64 one rounded)."""
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x = torch.zeros(b.shape, dtype=torch.float64, device=b.device)
        r = b.clone()
        q = torch.empty_like(b)
        rt = torch.empty_like(b)
        z = torch.empty_like(b)
        p = torch.empty_like(b)
        rr0 = self.dot(r, r)
        tol = max(abs_tol, rr0 * rel_tol)
        mode = "net" if use_net and self.net is not None else "vcycle"
        stats = {"net_iters": 0, "vcycle_iters": 0, "fallback_at": None, "replacements": 0, "handovers": 0}
        cap = max_iters if max_iters >= 0 else 20000
        it = 0
        rr = rr0
        true_rr = None

        def judge():
            """FP64 true residual of the FP64 x with the library's operator: the
            solve stops when it passes the test, else r takes it (at most
            max_replacements times)."""
            nonlocal rr, true_rr
            true_rr = self.true_residual(x, b, rt)
            if true_rr < tol or stats["replacements"] >= self.max_replacements:
                return True
            stats["replacements"] += 1
            r.copy_(rt)
            rr = true_rr
            return False

        def precond():
            nonlocal z
            if mode == "net":
                z = self.net_precondition(r, fluid)
                stats["net_iters"] += 1
            else:
                self.vcycle(r, z)
                stats["vcycle_iters"] += 1

        restart = True
        while it < cap:
            if mode == "vcycle" and handover:
                e, its = self.mg_solve(r, tol, max(cap - it, 0))
                x.add_(e)
                stats["vcycle_iters"] += its
                stats["handovers"] += 1
                it += its
                rr = self.mg.last_rTr
                if judge() or it >= cap:
                    break
                continue
            if restart:
                precond()
                p.copy_(z)
                zr = self.dot(z, r)
                restart = False
            self.apply_A(p, q)
            alpha = float(np.float32(zr / (self.dot(p, q) + eps)))   # the FP32 alpha of both updates
            x.add_(p, alpha=alpha)
            r.add_(q, alpha=-alpha)
            rr_new = self.dot(r, r)
            if not (rr_new == rr_new) or rr_new in (float("inf"),):
                raise FloatingPointError(f"non-finite residual at iteration {it + 1}")
            it += 1
            if mode == "net" and (rr_new > self.stagnation ** 2 * rr or stats["net_iters"] >= self.max_net_iters):
                mode = "vcycle"
                stats["fallback_at"] = it
                restart = True
            rr = rr_new
            if rr < tol:
                if judge():
                    break
                mode = "vcycle"
                restart = True
                continue
            if restart:
                continue
            precond()
            beta = -alpha * self.dot(z, q) / (zr + eps)
            zr = self.dot(z, r)
            torch.add(z, p, alpha=beta, out=p)   # p = z + beta p
        if true_rr is None:   # stopped by the iteration cap: judge once for the report
            true_rr = self.true_residual(x, b)
        x32 = x.float()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3
        stats.update(iterations=it, rTr=rr, true_rTr=true_rr, rTr0=rr0, ms=ms)
        self.last_x64 = x   # the FP64 solution (x32 is it rounded)
        self.last_b = b
        return x32, stats

    def fp32_true_residual(self):
        """||b - A x32||^2 of the last solve's FP32 solution (its FP64 x rounded),
        with the FP64 stencil: the accuracy of what the solve hands back."""
        return self.true_residual(self.last_x64.float().double(), self.last_b)
