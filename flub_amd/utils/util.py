"""Grid utilities on the projection path (utils/util.py:4-10 ``TexPair``,
:20-28 ``sample_3d``, :162-165 ``scale_by_3d``, :424-437 ``set_unknown`` of
the reference)."""
from __future__ import annotations

from .. import _lib
from ..launch import Kernel


class TexPair:
    """A current/next pair of buffers (utils/util.py:4-10)."""

    def __init__(self, cur, nxt):
        self.cur = cur
        self.nxt = nxt

    def swap(self):
        self.cur, self.nxt = self.nxt, self.cur


def _scale_by_3d(q, scale):
    _lib.call("flub_scale", q.f.ptr, float(scale), q.f.size, None)


scale_by_3d = Kernel("scale_by_3d", _scale_by_3d, lambda q, s: q.f.shape)


def _set_unknown(phi, tags):
    nx, ny, nz = phi.f.shape
    for d, a in ((1, tags.u), (2, tags.v), (3, tags.w)):
        want = (nx + (d == 1), ny + (d == 2), nz + (d == 3))
        if a.f.shape != want:
            raise ValueError("set_unknown: mark arrays do not match phi")
    _lib.call("flub_set_unknown", phi.f.ptr, tags.u.f.ptr, tags.v.f.ptr, tags.w.f.ptr, nx, ny, nz, None)


# utils/util.py:424-437
set_unknown = Kernel("set_unknown", _set_unknown, lambda phi, tags: phi.f.shape)


def sample_3d(qf, u: int, v: int, w: int):
    """Clamped fetch qf[clamp(u), clamp(v), clamp(w)] on a host (NumPy) array; the
    HIP kernels inline the same clamp in their stencils."""
    gx, gy, gz = qf.shape
    return qf[min(max(u, 0), gx - 1), min(max(v, 0), gy - 1), min(max(w, 0), gz - 1)]
