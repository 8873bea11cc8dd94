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
handle) with the absolute tolerance of the original test, and adds its
solution to x; with handover=False it continues this loop with the V-cycle
as preconditioner instead.  When the recursive residual passes the test, the
FP64 true residual of x is recomputed with the library's operator (MGPCG as
judge of the error); if it exceeds the tolerance by more than 4x it replaces
r and the solve continues on MGPCG.
"""
from __future__ import annotations

import ctypes
import math
import time

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

    @staticmethod
    def dot(a, b):
        return float(torch.dot(a.view(-1).double(), b.view(-1).double()))

    def _as_device(self, t):
        from ..device import DeviceArray
        return DeviceArray(tuple(t.shape), float, ptr=t.data_ptr(), owner=t)

    def mg_solve(self, r, tol, max_iters):
        """The library's fused MGPCG on residual r down to r.r < tol; returns
        (e, iterations) with e a view of the handle's x (valid until its next solve)."""
        import torch
        self.mg.init(self._as_device(r), 1.0)
        self.mg.mask_interior()
        its = self.mg.solve(max_iters, abs_tol=tol, rel_tol=0.0, smoother="phases")
        e = torch.empty_like(r)
        _lib.call("flub_memcpy_d2d", e.data_ptr(), self.mg.x.ptr, e.numel() * 4, None)
        return e, its

    def solve(self, b, fluid, max_iters=-1, rel_tol=1e-12, abs_tol=0.0, eps=1e-12, use_net=True, handover=True):
        """b: FP32 right-hand side (FLUID-masked); fluid: bool mask.  Returns
        (x, stats)."""
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x = torch.zeros_like(b)
        r = b.clone()
        q = torch.empty_like(b)
        ax = torch.empty_like(b)
        z = torch.empty_like(b)
        rr0 = self.dot(r, r)
        tol = max(abs_tol, rr0 * rel_tol)
        mode = "net" if use_net and self.net is not None else "vcycle"
        stats = {"net_iters": 0, "vcycle_iters": 0, "fallback_at": None, "replacements": 0, "handovers": 0}
        cap = max_iters if max_iters >= 0 else 20000
        it = 0
        rr = rr0

        def judge():
            """FP64 true residual with the library's operator; on drift (> 4x the
            tolerance) r takes it (at most max_replacements times).  True: stop."""
            nonlocal rr
            self.apply_A(x, ax)
            true_r = b.double() - ax.double()
            true_rr = float((true_r ** 2).sum())
            if true_rr <= 4.0 * tol or stats["replacements"] >= self.max_replacements:
                return True
            stats["replacements"] += 1
            r.copy_(true_r.float())
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
                p = z.clone()
                zr = self.dot(z, r)
                restart = False
            self.apply_A(p, q)
            alpha = zr / (self.dot(p, q) + eps)
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
            p = z + beta * p
        # judge: FP64 true residual with the library's operator
        self.apply_A(x, ax)
        true_rr = float(((b.double() - ax.double()) ** 2).sum())
        torch.cuda.synchronize()
        stats.update(iterations=it, rTr=rr, true_rTr=true_rr, rTr0=rr0, ms=(time.perf_counter() - t0) * 1e3)
        return x, stats
