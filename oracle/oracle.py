"""TEST INFRASTRUCTURE -- Python face of the CPU oracle (``liboracle.so``).

The oracle restates the reference Warp kernels of the projection path in C
(``flub_oracle.c``; every routine cites the reference file:line it follows).
It is the checker for tests/, the smoke check and the CPU-baseline leg of
bench.py, and nothing else: the shipped package never imports this module.

All arrays are NumPy, C-contiguous, k (z) fastest, as the reference lays out
its ``wp.array3d`` fields.  ``real`` selects the FP32 restatement or its FP64
twin.  Smoother ids: 0 = "default", 1 = "phases", 2 = "fraction".
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
SMOOTHERS = {"default": 0, "phases": 1, "fraction": 2}

_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_P = ctypes.c_void_p
_I = ctypes.c_int
_Z = ctypes.c_size_t
_D = ctypes.c_double


def _declare(L):
    for suf, R in (("_f32", ctypes.c_float), ("_f64", ctypes.c_double)):
        sig = {
            "orc_compute_divd": (None, [_P, _P, _I, _I, _I]),
            "orc_compute_Ap_mcadams": (None, [_P, _P, _P, _I, _I, _I]),
            "orc_compute_Ap_default": (None, [_P, _P, _I, _I, _I]),
            "orc_smooth_mcadams": (None, [_P, _P, _P, _P, _I, _I, _I, _I]),
            "orc_smooth_default": (None, [_P, _P, _I, _I, _I, _I]),
            "orc_restrict_mcadams_serial": (None, [_P, _P, _P, _P, _P, _I, _I, _I]),
            "orc_restrict_mcadams": (None, [_P, _P, _P, _P, _P, _I, _I, _I]),
            "orc_restrict_default": (None, [_P, _P, _P, _I, _I, _I]),
            "orc_prolongate": (None, [_P, _P, _I, _I, _I]),
            "orc_divergence_tag": (None, [_P] * 8 + [_I, _I, _I]),
            "orc_mask_with_tag": (None, [_P, _P, _Z]),
            "orc_scale": (None, [_P, R, _Z]),
            "orc_mask_interior": (None, [_P, _P, _Z]),
            "orc_subtract_gradient_tag": (None, [_P, _P, _P, _P, _I, _I, _I, _I]),
            "orc_zero_centered": (None, [_P, _P, _Z]),
            "orc_inner": (_D, [_P, _P, _Z, _I]),
            "orc_mgpcg_create": (_P, [_I, _I, _I, _I]),
            "orc_mgpcg_destroy": (None, [_P]),
            "orc_mgpcg_config": (None, [_P, _I, _I, _I, _I]),
            "orc_mgpcg_buf": (_P, [_P, _I, _I]),
            "orc_mgpcg_growth": (None, [_P, _I]),
            "orc_mgpcg_cycle": (None, [_P, ctypes.c_double, _I]),
            "orc_mgpcg_mean_f64": (None, [_P, _I]),
            "orc_mgpcg_tag": (_P, [_P, _I]),
            "orc_restrict_tag": (None, [_P, _P, _I, _I, _I]),
            "orc_mgpcg_set_tags": (None, [_P, _P]),
            "orc_mgpcg_init": (None, [_P, _P, R]),
            "orc_mgpcg_mask_interior": (None, [_P]),
            "orc_mgpcg_apply_preconditioner": (None, [_P, _I]),
            "orc_mgpcg_solve": (_I, [_P, _I, _D, _D, _D, _I, _I, ctypes.POINTER(_D)]),
            "orc_project": (_I, [_P] * 10 + [_D, _I, _D, _D, _D, _I, ctypes.POINTER(_D)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name + suf)
            f.restype = res
            f.argtypes = args
    L.orc_tag_dirichlet.argtypes = [_P, _P, _Z]
    L.orc_tag_neumann.argtypes = [_P, _P, _Z]
    L.orc_set_threads.argtypes = [_I]
    L.orc_set_threads.restype = _I
    _F = ctypes.c_float
    L.orc_sdf_cube.argtypes = [_P, _P, _P]
    L.orc_sdf_cube.restype = _F
    L.orc_sdf_atom_prims.argtypes = [_F, _F, _F, _P, _P, _I, _P]
    L.orc_sdf_atom_prims.restype = _F
    L.orc_sdf_argmin_prims.argtypes = [_F, _F, _F, _I, _P, _P, _P, _P, _F, _P]
    L.orc_sdf_argmin_prims.restype = _F
    L.orc_phi_prims.argtypes = [_P, _P, _I, _I, _I, _F, _F, _F, _F, _I, _P, _P, _P, _P]
    for name, n in (("orc_fraction_edge", 2), ("orc_fraction_triangle", 3), ("orc_fraction_quad", 5)):
        getattr(L, name).argtypes = [_F] * n
        getattr(L, name).restype = _F
    L.orc_phi_to_fraction.argtypes = [_P, _P, _I, _I, _I, _I]
    L.orc_update_argmin.argtypes = [_P, _P, _I, _I, _I, _P, _I, _I, _I, _I, _P, _P, _F, _P, _P, _F]
    L.orc_solid_vel.argtypes = [_P, _P, _I, _I, _I, _I, _F, _F, _F, _F, _P, _P, _P]
    L.orc_divergence_fractional.argtypes = [_P] * 10 + [_I, _I, _I]
    L.orc_set_unknown.argtypes = [_P, _P, _P, _P, _I, _I, _I]
    L.orc_sample_velocity.argtypes = [_P, _P, _P, _I, _I, _I, _P, _P, _I]
    L.orc_backtrace.argtypes = [_P, _P, _P, _I, _I, _I, _P, _I, _F, _F]
    L.orc_tricubic.argtypes = [_P, _I, _I, _I, _P, _P, _I]
    L.orc_sample_minmax.argtypes = [_P, _I, _I, _I, _P, _F, _P, _P, _P, _I]
    L.orc_advect_mc.argtypes = [_P, _P, _P, _I, _I, _I, _P, _P, _P, _I, _I, _I, _P, _F, _F]
    L.orc_advect_phi.argtypes = [_P, _P, _P, _I, _I, _I, _P, _P, _P, _F, _F]
    L.orc_advect_simple.argtypes = [_P, _P, _P, _I, _I, _I, _P, _P, _F, _F]
    L.orc_extrapolate_3d.argtypes = [_P, _P, _I, _I, _I]
    L.orc_flush_1.argtypes = [_P, _Z]


def set_threads(n: int) -> int:
    """Set the OpenMP thread count of the oracle (0 keeps it); returns it."""
    return lib().orc_set_threads(n)


def _ptr(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data


def _dt(real):
    return np.float32 if real in (np.float32, "f32", 32) else np.float64


def _suf(real):
    return "_f32" if _dt(real) == np.float32 else "_f64"


def _fn(name, real):
    return getattr(lib(), name + _suf(real))


def _arr(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


# ---------------------------------------------------------------- per kernel
def tag_dirichlet(phi):
    phi = _arr(phi, np.float32)
    tag = np.empty(phi.shape, np.int32)
    lib().orc_tag_dirichlet(_ptr(phi), _ptr(tag), phi.size)
    return tag


def tag_neumann(phi_rigid, tag):
    phi_rigid = _arr(phi_rigid, np.float32)
    tag = _arr(tag, np.int32).copy()
    lib().orc_tag_neumann(_ptr(phi_rigid), _ptr(tag), tag.size)
    return tag


def restrict_tag(tag):
    tag = _arr(tag, np.int32)
    cx, cy, cz = (s // 2 for s in tag.shape)
    out = np.empty((cx, cy, cz), np.int32)
    lib().orc_restrict_tag_f32(_ptr(tag), _ptr(out), cx, cy, cz)
    return out


def compute_divd(tag, real=np.float32):
    tag = _arr(tag, np.int32)
    out = np.empty(tag.shape, _dt(real))
    _fn("orc_compute_divd", real)(_ptr(tag), _ptr(out), *tag.shape)
    return out


def compute_Ap(p, tag=None, real=np.float32):
    p = _arr(p, _dt(real))
    Ap = np.empty_like(p)
    if tag is None:
        _fn("orc_compute_Ap_default", real)(_ptr(Ap), _ptr(p), *p.shape)
    else:
        tag = _arr(tag, np.int32)
        _fn("orc_compute_Ap_mcadams", real)(_ptr(Ap), _ptr(p), _ptr(tag), *p.shape)
    return Ap


def smooth(z, r, tag, divd, phase, real=np.float32):
    """One _smooth_mcadams launch (tag given) or _smooth (tag None); returns new z."""
    z = _arr(z, _dt(real)).copy()
    r = _arr(r, _dt(real))
    if tag is None:
        _fn("orc_smooth_default", real)(_ptr(z), _ptr(r), *z.shape, int(phase))
    else:
        tag = _arr(tag, np.int32)
        divd = _arr(divd, _dt(real))
        _fn("orc_smooth_mcadams", real)(_ptr(z), _ptr(r), _ptr(tag), _ptr(divd), *z.shape, int(phase))
    return z


def restrict(r, z, tag=None, tag_c=None, real=np.float32, serial=False):
    r = _arr(r, _dt(real))
    z = _arr(z, _dt(real))
    out = np.zeros(tuple(s // 2 for s in r.shape), _dt(real))
    if tag is None:
        _fn("orc_restrict_default", real)(_ptr(r), _ptr(z), _ptr(out), *r.shape)
    else:
        tag = _arr(tag, np.int32)
        tag_c = _arr(tag_c, np.int32)
        name = "orc_restrict_mcadams_serial" if serial else "orc_restrict_mcadams"
        _fn(name, real)(_ptr(r), _ptr(z), _ptr(tag), _ptr(out), _ptr(tag_c), *r.shape)
    return out


def prolongate(z, zc, real=np.float32):
    z = _arr(z, _dt(real)).copy()
    zc = _arr(zc, _dt(real))
    _fn("orc_prolongate", real)(_ptr(z), _ptr(zc), *z.shape)
    return z


def divergence_tag(u, v, w, ur, vr, wr, tag, real=np.float32):
    dt = _dt(real)
    u, v, w, ur, vr, wr = (_arr(a, dt) for a in (u, v, w, ur, vr, wr))
    tag = _arr(tag, np.int32)
    div = np.empty(tag.shape, dt)
    _fn("orc_divergence_tag", real)(_ptr(u), _ptr(v), _ptr(w), _ptr(div), _ptr(ur), _ptr(vr),
                                    _ptr(wr), _ptr(tag), *tag.shape)
    return div


def mask_with_tag(div, tag, real=np.float32):
    div = _arr(div, _dt(real)).copy()
    tag = _arr(tag, np.int32)
    _fn("orc_mask_with_tag", real)(_ptr(div), _ptr(tag), div.size)
    return div


def scale(q, s, real=np.float32):
    q = _arr(q, _dt(real)).copy()
    _fn("orc_scale", real)(_ptr(q), s, q.size)
    return q


def subtract_gradient_tag(uf, u_rigid, p, tag, d, real=np.float32):
    dt = _dt(real)
    uf = _arr(uf, dt).copy()
    u_rigid = _arr(u_rigid, dt)
    p = _arr(p, dt)
    tag = _arr(tag, np.int32)
    _fn("orc_subtract_gradient_tag", real)(_ptr(uf), _ptr(u_rigid), _ptr(p), _ptr(tag), *tag.shape, int(d))
    return uf


def zero_centered(q, tag, real=np.float32):
    q = _arr(q, _dt(real)).copy()
    tag = _arr(tag, np.int32)
    _fn("orc_zero_centered", real)(_ptr(q), _ptr(tag), q.size)
    return q


def inner(a, b, real=np.float32, f64=True):
    a = _arr(a, _dt(real))
    b = _arr(b, _dt(real))
    return _fn("orc_inner", real)(_ptr(a), _ptr(b), a.size, int(f64))


# ---------------------------------------------------------------- MGPCG
class MGPCG:
    """Oracle twin of multigrid/mgpcg.py:234-490 (same levels and control flow)."""

    def __init__(self, shape, n_mg_levels, real=np.float32, pre_and_post_smoothing=6,
                 bottom_smoothing=6, dot_f64=True, parallel_dot=False):
        self.real = _dt(real)
        self.shape = tuple(int(s) for s in shape)
        self.L = int(n_mg_levels)
        self._h = _fn("orc_mgpcg_create", self.real)(*self.shape, self.L)
        if not self._h:
            raise ValueError("bad level count")
        _fn("orc_mgpcg_config", self.real)(self._h, int(pre_and_post_smoothing), int(bottom_smoothing),
                                           int(dot_f64), int(parallel_dot))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _fn("orc_mgpcg_destroy", self.real)(h)
            self._h = None

    def set_mean_f64(self, flag=True):
        """zero_centered inside solve sums its mean in double (the HIP path's
        accumulation) instead of the reference's serial FP32 sum."""
        _fn("orc_mgpcg_mean_f64", self.real)(self._h, int(bool(flag)))

    def set_growth(self, growth):
        """Smoothing count growth per level (1: the reference's << l; 0: constant)."""
        _fn("orc_mgpcg_growth", self.real)(self._h, int(growth))

    def set_cycle(self, coarse_scale=1.0, sweep_cap=0):
        """The V-cycle variant without a reference counterpart: restricted
        residuals times coarse_scale, at most sweep_cap sweeps per level
        (1.0 and 0: the reference)."""
        _fn("orc_mgpcg_cycle", self.real)(self._h, float(coarse_scale), int(sweep_cap))

    def level_shape(self, l):
        return tuple(s >> l for s in self.shape)

    def _view(self, which, level, dtype=None):
        shp = self.level_shape(level)
        dtype = dtype or self.real
        if which == "tag":
            ptr = _fn("orc_mgpcg_tag", self.real)(self._h, level)
            dtype = np.int32
        else:
            code = {"x": 0, "p": 1, "Ap": 2, "r": 3, "z": 4, "divd": 5}[which]
            ptr = _fn("orc_mgpcg_buf", self.real)(self._h, code, level)
        n = int(np.prod(shp))
        buf = (ctypes.c_byte * (n * np.dtype(dtype).itemsize)).from_address(ptr)
        return np.frombuffer(buf, dtype=dtype).reshape(shp)

    def get(self, which, level=0):
        return self._view(which, level).copy()

    def put(self, which, arr, level=0):
        self._view(which, level)[...] = arr

    def set_tags(self, tag=None):
        if tag is None:
            _fn("orc_mgpcg_set_tags", self.real)(self._h, None)
        else:
            tag = _arr(tag, np.int32)
            _fn("orc_mgpcg_set_tags", self.real)(self._h, _ptr(tag))

    def init(self, r, k):
        r = _arr(r, self.real)
        _fn("orc_mgpcg_init", self.real)(self._h, _ptr(r), k)

    def mask_interior(self):
        _fn("orc_mgpcg_mask_interior", self.real)(self._h)

    def apply_preconditioner(self, smoother="phases"):
        _fn("orc_mgpcg_apply_preconditioner", self.real)(self._h, SMOOTHERS[smoother])

    def solve(self, max_iters=1, eps=1e-12, abs_tol=1e-12, rel_tol=1e-12, smoother="default", verbose=False):
        rtr = ctypes.c_double(0.0)
        it = _fn("orc_mgpcg_solve", self.real)(self._h, int(max_iters), eps, abs_tol, rel_tol,
                                              SMOOTHERS[smoother], int(verbose), ctypes.byref(rtr))
        return it, rtr.value

    def project(self, u, v, w, ur, vr, wr, tag, h, max_iters=-1, abs_tol=0.0, rel_tol=1e-12,
                smoother="phases", eps=None):
        """example_fluid.py:120-147 (+ water_3d.py:58-68). Returns (u, v, w, p0, div, iters, rTr)."""
        dt = self.real
        u, v, w = (_arr(a, dt).copy() for a in (u, v, w))
        ur, vr, wr = (_arr(a, dt) for a in (ur, vr, wr))
        tag = _arr(tag, np.int32)
        div = np.empty(self.shape, dt)
        p0 = np.empty(self.shape, dt)
        rtr = ctypes.c_double(0.0)
        if eps is None:
            eps = 1e-12 if dt == np.float32 else 0.0
        it = _fn("orc_project", dt)(self._h, _ptr(u), _ptr(v), _ptr(w), _ptr(ur), _ptr(vr), _ptr(wr),
                                    _ptr(tag), _ptr(div), _ptr(p0), float(h), int(max_iters), float(eps),
                                    float(abs_tol), float(rel_tol), SMOOTHERS[smoother], ctypes.byref(rtr))
        return u, v, w, p0, div, it, rtr.value


# ------------------------------------------------ solid geometry (FP32 only)
STAGGER = {0: (0.5, 0.5, 0.5), 1: (0.0, 0.5, 0.5), 2: (0.5, 0.0, 0.5), 3: (0.5, 0.5, 0.0)}


class Rigids:
    """Rigid bodies as arrays: x (n,3), q (n,4) as (x, y, z, w), gtype (n,),
    radius (n,3), v (n,3), omega (n,3); all float32 / int32."""

    def __init__(self, x, q, gtype, radius, v=None, omega=None):
        n = len(gtype)
        self.x = _arr(np.reshape(x, (n, 3)), np.float32)
        self.q = _arr(np.reshape(q, (n, 4)), np.float32)
        self.gtype = _arr(gtype, np.int32)
        self.radius = _arr(np.reshape(radius, (n, 3)), np.float32)
        self.v = _arr(np.zeros((n, 3)) if v is None else np.reshape(v, (n, 3)), np.float32)
        self.omega = _arr(np.zeros((n, 3)) if omega is None else np.reshape(omega, (n, 3)), np.float32)
        self.n = n


def sdf_cube(p, radius, q):
    """geometry.py:17-32 at one point."""
    p, radius, q = (_arr(a, np.float32) for a in (p, radius, q))
    return float(lib().orc_sdf_cube(_ptr(p), _ptr(radius), _ptr(q)))


def sdf_atom_prims(p, rb: Rigids, b: int):
    """geometry.py:82-113 for body b at point p."""
    return float(lib().orc_sdf_atom_prims(float(p[0]), float(p[1]), float(p[2]), _ptr(rb.x[b]), _ptr(rb.q[b]),
                                          int(rb.gtype[b]), _ptr(rb.radius[b])))


def phi_prims(shape, stagger, h, rb: Rigids, owner=True):
    """geometry.py:225-239 over an array of ``shape``; returns (phi, owner or None)."""
    phi = np.empty(shape, np.float32)
    own = np.empty(shape, np.int32) if owner else None
    lib().orc_phi_prims(_ptr(phi), _ptr(own) if owner else None, *shape, *(float(s) for s in stagger), float(h),
                        rb.n, _ptr(rb.x), _ptr(rb.q), _ptr(rb.gtype), _ptr(rb.radius))
    return phi, own


def update_argmin(phi, owner, levelset, body_id, x, q, h_levelset, stagger, ijk, h):
    """geometry.py:365-388 on copies of phi and owner (owner may be None);
    returns (phi, owner)."""
    phi = _arr(phi, np.float32).copy()
    own = None if owner is None else _arr(owner, np.int32).copy()
    ls = _arr(levelset, np.float32)
    x, q, st, ijk = (_arr(np.reshape(a, -1), np.float32) for a in (x, q, stagger, ijk))
    lib().orc_update_argmin(_ptr(phi), None if own is None else _ptr(own), *phi.shape, _ptr(ls), *ls.shape,
                            int(body_id), _ptr(x), _ptr(q), float(h_levelset), _ptr(st), _ptr(ijk), float(h))
    return phi, own


def fraction_edge(yl, yr):
    return float(lib().orc_fraction_edge(yl, yr))


def fraction_triangle(ya, yb, yc):
    return float(lib().orc_fraction_triangle(ya, yb, yc))


def fraction_quad(ya, yb, yc, yd, ycen):
    return float(lib().orc_fraction_quad(ya, yb, yc, yd, ycen))


def phi_to_fraction(phi_corner, frac_centre, dim):
    """geometry.py:264-290; frac_centre holds the face-centre phi; returns the fractions."""
    phi_corner = _arr(phi_corner, np.float32)
    frac = _arr(frac_centre, np.float32).copy()
    nx, ny, nz = (s - 1 for s in phi_corner.shape)
    lib().orc_phi_to_fraction(_ptr(phi_corner), _ptr(frac), nx, ny, nz, int(dim))
    return frac


def face_fractions(shape, h, rb: Rigids):
    """geometry.py:440-462 precompute_fij_with_cache for primitive bodies:
    returns ((fu, fv, fw), (owner_u, owner_v, owner_w))."""
    nx, ny, nz = shape
    corner, _ = phi_prims((nx + 1, ny + 1, nz + 1), (0.0, 0.0, 0.0), h, rb, owner=False)
    fr, ow = [], []
    for d, fs in ((1, (nx + 1, ny, nz)), (2, (nx, ny + 1, nz)), (3, (nx, ny, nz + 1))):
        cen, own = phi_prims(fs, STAGGER[d], h, rb)
        fr.append(phi_to_fraction(corner, cen, d))
        ow.append(own)
    return tuple(fr), tuple(ow)


def solid_vel(owner, d, h0, rb: Rigids):
    """rigid_coupling.py:51-70 precompute_solid_vel for axis d over the face array owner."""
    owner = _arr(owner, np.int32)
    vel = np.empty(owner.shape, np.float32)
    st = STAGGER[d]
    lib().orc_solid_vel(_ptr(vel), _ptr(owner), *owner.shape, int(d), *(float(s) for s in st), float(h0),
                        _ptr(rb.x), _ptr(rb.v), _ptr(rb.omega))
    return vel


def divergence_fractional(u, v, w, fu, fv, fw, ur, vr, wr):
    """divergence.py:5-41 divergence_fractional_3d; returns div."""
    arrs = [_arr(a, np.float32) for a in (u, v, w, fu, fv, fw, ur, vr, wr)]
    u, v, w, fu, fv, fw, ur, vr, wr = arrs
    nx, ny, nz = v.shape[0], u.shape[1], u.shape[2]
    div = np.empty((nx, ny, nz), np.float32)
    lib().orc_divergence_fractional(_ptr(u), _ptr(v), _ptr(w), _ptr(div), _ptr(fu), _ptr(fv), _ptr(fw), _ptr(ur),
                                    _ptr(vr), _ptr(wr), nx, ny, nz)
    return div


# ------------------------------------------------------------ extrapolation
def set_unknown(phi, tu, tv, tw):
    """utils/util.py:424-437; returns new (tu, tv, tw) (last face layers kept)."""
    phi = _arr(phi, np.float32)
    tu, tv, tw = (_arr(t, np.int32).copy() for t in (tu, tv, tw))
    lib().orc_set_unknown(_ptr(phi), _ptr(tu), _ptr(tv), _ptr(tw), *phi.shape)
    return tu, tv, tw


def extrapolate_3d(q, unknown):
    """extrapolate.py:9-31 on one face array; returns new (q, unknown)."""
    q = _arr(q, np.float32).copy()
    unknown = _arr(unknown, np.int32).copy()
    lib().orc_extrapolate_3d(_ptr(q), _ptr(unknown), *q.shape)
    return q, unknown


def flush_1(tag):
    tag = _arr(tag, np.int32).copy()
    lib().orc_flush_1(_ptr(tag), tag.size)
    return tag


def extrapolate_by_1(vel, tags):
    """extrapolate.py:40-47 on (u, v, w) and their tags; returns (vel, tags)."""
    out_v, out_t = [], []
    for q, t in zip(vel, tags):
        q, t = extrapolate_3d(q, t)
        out_v.append(q)
        out_t.append(t)
    return tuple(out_v), tuple(flush_1(t) for t in out_t)


# ---------------------------------------------------------------- advection
def _mac3(u, v, w):
    u, v, w = (_arr(a, np.float32) for a in (u, v, w))
    nx, ny, nz = v.shape[0], u.shape[1], u.shape[2]
    return u, v, w, nx, ny, nz


def sample_velocity(u, v, w, pts):
    """utils/util.py:33-86 at points (n, 3) in grid units; returns (n, 3)."""
    u, v, w, nx, ny, nz = _mac3(u, v, w)
    pts = _arr(np.reshape(pts, (-1, 3)), np.float32)
    out = np.empty_like(pts)
    lib().orc_sample_velocity(_ptr(u), _ptr(v), _ptr(w), nx, ny, nz, _ptr(pts), _ptr(out), len(pts))
    return out


def backtrace(u, v, w, pts, dt, h):
    """advect.py:11-22 (RK3) of points (n, 3) in grid units; returns (n, 3)."""
    u, v, w, nx, ny, nz = _mac3(u, v, w)
    pts = _arr(np.reshape(pts, (-1, 3)), np.float32).copy()
    lib().orc_backtrace(_ptr(u), _ptr(v), _ptr(w), nx, ny, nz, _ptr(pts), len(pts), float(dt), float(h))
    return pts


def tricubic(q, str_pts):
    """utils/util.py:130-159 _tricubic at (s, t, r) points (n, 3), index units."""
    q = _arr(q, np.float32)
    pts = _arr(np.reshape(str_pts, (-1, 3)), np.float32)
    out = np.empty(len(pts), np.float32)
    lib().orc_tricubic(_ptr(q), *q.shape, _ptr(pts), _ptr(out), len(pts))
    return out


def sample_minmax(q, pts, h, stagger):
    """advect.py:114-131; returns (min, max) arrays."""
    q = _arr(q, np.float32)
    pts = _arr(np.reshape(pts, (-1, 3)), np.float32)
    st = _arr(stagger, np.float32)
    mn = np.empty(len(pts), np.float32)
    mx = np.empty(len(pts), np.float32)
    lib().orc_sample_minmax(_ptr(q), *q.shape, _ptr(pts), float(h), _ptr(st), _ptr(mn), _ptr(mx), len(pts))
    return mn, mx


def advect_mc(u, v, w, q_in, stagger, dt, h0):
    """advect.py:58-112,133-137 (the three launches); returns q_out."""
    u, v, w, nx, ny, nz = _mac3(u, v, w)
    q_in = _arr(q_in, np.float32)
    q_out = np.empty_like(q_in)
    q_tmp = np.empty_like(q_in)
    st = _arr(stagger, np.float32)
    lib().orc_advect_mc(_ptr(u), _ptr(v), _ptr(w), nx, ny, nz, _ptr(q_in), _ptr(q_out), _ptr(q_tmp), *q_in.shape,
                        _ptr(st), float(dt), float(h0))
    return q_out


def advect_phi(u, v, w, phi, dt, h0, stagger=(0.5, 0.5, 0.5)):
    """advect.py:38-56; returns new phi."""
    u, v, w, nx, ny, nz = _mac3(u, v, w)
    phi = _arr(phi, np.float32)
    out = np.empty_like(phi)
    st = _arr(stagger, np.float32)
    lib().orc_advect_phi(_ptr(u), _ptr(v), _ptr(w), nx, ny, nz, _ptr(phi), _ptr(out), _ptr(st), float(dt), float(h0))
    return out


def advect_simple(u, v, w, q, dt, h):
    """advect.py:25-35; returns new q."""
    u, v, w, nx, ny, nz = _mac3(u, v, w)
    q = _arr(q, np.float32)
    out = np.empty_like(q)
    lib().orc_advect_simple(_ptr(u), _ptr(v), _ptr(w), nx, ny, nz, _ptr(q), _ptr(out), float(dt), float(h))
    return out
