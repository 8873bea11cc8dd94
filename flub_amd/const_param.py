"""Grid configuration and MAC-grid containers (const_param.py:34-60,80-150 of
the reference): ``Field``/``IntField`` (f, stagger, d), ``MACGrid``/``MACGridInt``
(u, v, w), ``make_field``, ``make_mac_grid``, ``field_like``, ``get_uf``.

The reference fixes the grid in module globals; here they are set with
``set_grid(gx, gy, gz)`` (defaults: the reference's 256 x 64 x 256).
"""
from __future__ import annotations

from dataclasses import dataclass, field as _dc_field

from . import device as wp_dev

gridx, gridy, gridz = 256, 64, 256
res = 256
shape = (gridx, gridy, gridz)
shapeu = (gridx + 1, gridy, gridz)
shapev = (gridx, gridy + 1, gridz)
shapew = (gridx, gridy, gridz + 1)
n_unknowns = gridx * gridy * gridz
h = 1.0 / float(gridx)
h0 = h
dt = 1e-2


def set_grid(gx: int, gy: int, gz: int) -> None:
    """Reconfigure the module-level grid (the reference edits these globals)."""
    global gridx, gridy, gridz, res, shape, shapeu, shapev, shapew, n_unknowns, h, h0
    gridx, gridy, gridz = int(gx), int(gy), int(gz)
    res = gridx
    shape = (gridx, gridy, gridz)
    shapeu = (gridx + 1, gridy, gridz)
    shapev = (gridx, gridy + 1, gridz)
    shapew = (gridx, gridy, gridz + 1)
    n_unknowns = gridx * gridy * gridz
    h = 1.0 / float(gridx)
    h0 = h


@dataclass
class Field:
    f: wp_dev.DeviceArray = None
    stagger: tuple = (0.5, 0.5, 0.5)
    d: int = 0


@dataclass
class IntField:
    f: wp_dev.DeviceArray = None
    stagger: tuple = (0.5, 0.5, 0.5)
    d: int = 0


@dataclass
class MACGrid:
    u: Field = _dc_field(default_factory=Field)
    v: Field = _dc_field(default_factory=Field)
    w: Field = _dc_field(default_factory=Field)


@dataclass
class MACGridInt:
    u: IntField = _dc_field(default_factory=IntField)
    v: IntField = _dc_field(default_factory=IntField)
    w: IntField = _dc_field(default_factory=IntField)


def get_uf(vels, d: int):
    """d = 1, 2, 3 for u, v, w respectively (const_param.py:104-110)."""
    return vels.u if d == 1 else (vels.v if d == 2 else vels.w)


_STAGGER = {0: (0.5, 0.5, 0.5), 1: (0.0, 0.5, 0.5), 2: (0.5, 0.0, 0.5), 3: (0.5, 0.5, 0.0)}


def field_shape(dim: int, grid=None):
    gx, gy, gz = grid if grid is not None else shape
    return {0: (gx, gy, gz), 1: (gx + 1, gy, gz), 2: (gx, gy + 1, gz), 3: (gx, gy, gz + 1)}[dim]


def make_field(dim, dtype=float, grid=None):
    """const_param.py:114-133: a zeroed centred (dim 0) or face (1, 2, 3) field."""
    cls = Field if dtype in (float, "float32") else IntField
    return cls(f=wp_dev.zeros(field_shape(dim, grid), dtype), stagger=_STAGGER[dim], d=dim)


def make_mac_grid(dtype=float, grid=None):
    """const_param.py:135-143."""
    if dtype in (float, "float32"):
        return MACGrid(make_field(1, dtype, grid), make_field(2, dtype, grid), make_field(3, dtype, grid))
    return MACGridInt(make_field(1, dtype, grid), make_field(2, dtype, grid), make_field(3, dtype, grid))


def field_like(fld):
    """const_param.py:145-150."""
    return type(fld)(f=wp_dev.zeros_like(fld.f), stagger=fld.stagger, d=fld.d)
