"""Cell classification (geometry.py:335-344 ``_tag_dirichlet`` and :390-397
``_tag_neumann`` of the reference) on the HIP C-ABI, in the order
example_fluid.py:111-114 applies them: water phi -> AIR/FLUID, then rigid
phi < 0 -> SOLID."""
from __future__ import annotations

from . import _lib
from .launch import Kernel


def _tag_dirichlet(tag, phi):
    if tag.f.shape != phi.f.shape:
        raise ValueError("tag_dirichlet: shape mismatch")
    _lib.call("flub_tag_dirichlet", phi.f.ptr, tag.f.ptr, tag.f.size, None)


def _tag_neumann(tag, phi):
    if tag.f.shape != phi.f.shape:
        raise ValueError("tag_neumann: shape mismatch")
    _lib.call("flub_tag_neumann", phi.f.ptr, tag.f.ptr, tag.f.size, None)


tag_dirichlet_kernel = Kernel("_tag_dirichlet", _tag_dirichlet, lambda tag, phi: tag.f.shape)
tag_neumann_kernel = Kernel("_tag_neumann", _tag_neumann, lambda tag, phi: tag.f.shape)


def tag_dirichlet(tag, phi):
    _tag_dirichlet(tag, phi)


def tag_neumann(tag, phi_rigid):
    _tag_neumann(tag, phi_rigid)
