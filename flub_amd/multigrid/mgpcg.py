"""MGPCG pressure solver -- drop-in for the reference's ``multigrid/mgpcg.py``
class ``MGPCG`` (mgpcg.py:234-490), running on the HIP C-ABI (libflub.so).

Same constructor, attributes and methods:
``MGPCG(dim=3, shape=(256,256,256), n_mg_levels=6, real=float)``;
``set_tags(tag=None)``, ``init(r, k)``, ``mask_interior()``, ``zero_centered()``,
``apply_preconditioner(smoother)``, ``graph_precondition(smoother)``,
``update_p/r/x()``, ``update_alpha/beta(eps)``, ``solve(max_iters=1, eps=1e-12,
abs_tol=1e-12, rel_tol=1e-12, verbose=False, smoother="default")``,
``analyze(tag0)``; attributes ``x, p, Ap, r[], z[], tag[], divd[], alpha, beta,
sum, old_zTr, n_mg_levels, pre_and_post_smoothing, bottom_smoothing, shape``.

Differences, all additive: ``precision`` selects an FP16/BF16 V-cycle inside
flexible CG ("fp32" keeps the reference algorithm); ``solve`` returns the
completed-iteration count; ``tag[l]`` holds one byte per cell; ``divd[l]`` is
computed from the tags on access (the solver never stores it); the scalars
are FP64.  The first ``graph_precondition`` call runs the V-cycle (the
reference's first call only captured it) and ``beta`` starts at 0.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _lib
from ..device import DeviceArray

SMOOTHERS = {"default": 0, "phases": 1, "fraction": 2}
PRECISIONS = {"fp32": 0, "fp16": 1, "bf16": 2}
_SCALARS = {"alpha": 0, "beta": 1, "sum": 2, "old_zTr": 3, "rTr": 4}
_BUFS = {"x": 0, "p": 1, "Ap": 2, "r": 3, "z": 4, "tag": 5, "scalars": 6}


class _DivdLevels:
    """divd[l] of the reference (mgpcg.py:51-58), computed from the packed tags."""

    def __init__(self, mg):
        self._mg = mg
        self._cache = {}

    def __len__(self):
        return self._mg.n_mg_levels

    def __getitem__(self, l):
        l = range(self._mg.n_mg_levels)[l]
        if l not in self._cache:
            self._cache[l] = DeviceArray(self._mg.level_shape(l), np.float32)
        out = self._cache[l]
        _lib.call("flub_mgpcg_compute_divd", self._mg._h, l, out.ptr)
        return out


class MGPCG:
    def __init__(self, dim=3, shape=(256, 256, 256), n_mg_levels=6, real=float, precision="fp32",
                 stream=None):
        if dim != 3:
            raise ValueError("only dim=3 is supported (as in the reference)")
        self.dim = dim
        self.real = real
        self.shape = tuple(int(s) for s in shape)
        self.n_mg_levels = int(n_mg_levels)
        self.precision = precision
        h = ctypes.c_void_p()
        _lib.call("flub_mgpcg_create", ctypes.byref(h), *self.shape, self.n_mg_levels,
                  PRECISIONS[precision], stream)
        self._h = h.value
        self._pre_post = 6
        self._bottom = 6
        self._coarse_scale = 1.0
        self._sweep_cap = 0
        self.r = [self._view("r", l, np.float32) for l in range(self.n_mg_levels)]
        self.z = [self._view("z", l, np.float32) for l in range(self.n_mg_levels)]
        self.tag = [self._view("tag", l, np.uint8) for l in range(self.n_mg_levels)]
        self.divd = _DivdLevels(self)
        self.x = self._view("x", 0, np.float32)
        self.p = self._view("p", 0, np.float32)
        self.Ap = self._view("Ap", 0, np.float32)
        sc = ctypes.c_void_p()
        _lib.call("flub_mgpcg_buffer", self._h, _BUFS["scalars"], 0, ctypes.byref(sc))
        for name, i in _SCALARS.items():
            setattr(self, name if name != "rTr" else "rTr_dev",
                    DeviceArray((1,), np.float64, ptr=sc.value + 8 * i, owner=self))
        self.graphs = {"default": None, "phases": None, "fraction": None}
        self.last_iterations = None
        self.last_rTr = None

    # ------------------------------------------------------------ plumbing
    def level_shape(self, l):
        s = (ctypes.c_int * 3)()
        _lib.call("flub_mgpcg_level_shape", self._h, int(l), s)
        return tuple(s)

    def _view(self, which, level, dtype):
        p = ctypes.c_void_p()
        _lib.call("flub_mgpcg_buffer", self._h, _BUFS[which], level, ctypes.byref(p))
        return DeviceArray(self.level_shape(level), dtype, ptr=p.value, owner=self)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.lib().flub_mgpcg_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    @property
    def pre_and_post_smoothing(self):
        return self._pre_post

    @pre_and_post_smoothing.setter
    def pre_and_post_smoothing(self, v):
        self._pre_post = int(v)
        _lib.call("flub_mgpcg_set_smoothing", self._h, self._pre_post, self._bottom)

    @property
    def bottom_smoothing(self):
        return self._bottom

    @bottom_smoothing.setter
    def bottom_smoothing(self, v):
        self._bottom = int(v)
        _lib.call("flub_mgpcg_set_smoothing", self._h, self._pre_post, self._bottom)

    @property
    def coarse_scale(self):
        """Scale of the coarse-grid correction (no reference counterpart): each
        restriction stores coarse_scale times the children's residual sum.
        1.0 (the default) is the reference's V-cycle, bitwise."""
        return self._coarse_scale

    @coarse_scale.setter
    def coarse_scale(self, v):
        _lib.call("flub_mgpcg_set_coarse_scale", self._h, float(v))
        self._coarse_scale = float(v)

    @property
    def sweep_cap(self):
        """At most this many pre/post red-black iterations on any level (no
        reference counterpart; 0, the default: the reference's doubling)."""
        return self._sweep_cap

    @sweep_cap.setter
    def sweep_cap(self, v):
        _lib.call("flub_mgpcg_set_sweep_cap", self._h, int(v))
        self._sweep_cap = int(v)

    def scalars(self):
        """(alpha, beta, sum, old_zTr, rTr) from the device, as floats."""
        out = (ctypes.c_double * 5)()
        _lib.call("flub_mgpcg_scalars", self._h, out)
        return tuple(out)

    # ------------------------------------------------------------ reference API
    def zero_centered(self):
        _lib.call("flub_mgpcg_zero_centered", self._h)

    def init(self, r, k):
        if r.shape != self.shape:
            raise ValueError(f"init: r has shape {r.shape}, expected {self.shape}")
        _lib.call("flub_mgpcg_init", self._h, r.ptr, float(k))

    def graph_precondition(self, smoother="default"):
        assert smoother in SMOOTHERS
        _lib.call("flub_mgpcg_graph_precondition", self._h, SMOOTHERS[smoother])
        self.graphs[smoother] = True

    def apply_preconditioner(self, smoother="default"):
        _lib.call("flub_mgpcg_apply_preconditioner", self._h, SMOOTHERS[smoother])

    def compute_Ap(self, smoother="phases"):
        """compute_Ap_mcadams / compute_Ap on p (mgpcg.py:408-414); also leaves p.Ap in ``sum``."""
        _lib.call("flub_mgpcg_compute_Ap", self._h, SMOOTHERS[smoother])

    def update_p(self):
        _lib.call("flub_mgpcg_update_p", self._h)

    def update_r(self):
        _lib.call("flub_mgpcg_update_r", self._h)

    def update_x(self):
        _lib.call("flub_mgpcg_update_x", self._h)

    def set_tags(self, tag=None):
        if tag is None:
            _lib.call("flub_mgpcg_set_tags", self._h, None)
            return
        if tag.shape != self.shape or tag.dtype != np.int32:
            raise ValueError(f"set_tags: expected int32 {self.shape}, got {tag.dtype} {tag.shape}")
        _lib.call("flub_mgpcg_set_tags", self._h, tag.ptr)

    def update_alpha(self, eps):
        _lib.call("flub_mgpcg_update_alpha", self._h, float(eps))

    def update_beta(self, eps):
        _lib.call("flub_mgpcg_update_beta", self._h, float(eps))

    def inner(self, a: str, b: str, out: str = "sum"):
        """array_inner(self.<a>, self.<b>, self.<out>) for a, b in x, p, Ap, r, z."""
        ids = {"x": 0, "p": 1, "Ap": 2, "r": 3, "z": 4}
        _lib.call("flub_mgpcg_inner", self._h, ids[a], ids[b], _SCALARS[out])

    def solve(self, max_iters=1, eps=1e-12, abs_tol=1e-12, rel_tol=1e-12, verbose=False,
              smoother="default"):
        """mgpcg.py:378-462.  Returns the completed-iteration count."""
        it = ctypes.c_int(0)
        rtr = ctypes.c_double(0.0)
        _lib.call("flub_mgpcg_solve", self._h, int(max_iters), float(eps), float(abs_tol),
                  float(rel_tol), SMOOTHERS[smoother], int(bool(verbose)), ctypes.byref(it),
                  ctypes.byref(rtr), allow=(_lib.ENOTCONV,))
        self.last_iterations = it.value
        self.last_rTr = rtr.value
        return it.value

    def mask_interior(self):
        _lib.call("flub_mgpcg_mask_interior", self._h)

    def analyze(self, tag0):
        """mgpcg.py:467-490: retag small enclosed fluid pockets (no AIR cell, 2..49
        cells, face-connected) as SOLID.  Host-side diagnostic on a NumPy tag array
        (scipy.ndimage.label with face connectivity in place of skimage's label)."""
        from scipy import ndimage

        tag0 = np.array(tag0, copy=True)
        open_cells = tag0 != 2
        labels, num = ndimage.label(open_cells, structure=ndimage.generate_binary_structure(3, 1))
        print("num_features = ", num)
        for f in range(num):
            idx = np.argwhere(labels == f + 1)
            print(f"component {f} length = {len(idx)}")
            if 1 < len(idx) < 50:
                if not np.any(tag0[tuple(idx.T)] == 0):
                    print("trouble component found")
                    tag0[tuple(idx.T)] = 2
        return tag0
