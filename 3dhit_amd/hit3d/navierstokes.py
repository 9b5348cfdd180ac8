"""Rotational-form Navier–Stokes RHS (navierstokes.py of hit3d):

    du/dt = P(k) [ (u x w)^ + f ] - nu k^2 u

`make_ns_rhs(grid)` returns a hashable callable ``rhs(u_hat, params, t)``.
Each call is one pass sequence of hand-written HIP kernels (x/y/z pencil
FFTs with the curl, the cross product, dealiasing, projection, forcing and
viscous term fused in) through the C ABI function `hit3d_rhs`.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import torch

from . import _lib
from .forcing import NoForcing
from .grid import check_spectral, stream

__all__ = ["NSParams", "make_ns_rhs", "linear_operator", "is_dealiased"]


@dataclass(frozen=True)
class NSParams:
    """navierstokes.py:40-59: viscosity + forcing."""

    nu: float
    forcing: object = field(default_factory=NoForcing)

    @property
    def label(self):
        return f"NavierStokes_nu{float(self.nu):g}_{self.forcing.label}"

    def __repr__(self):
        return f"NavierStokes(nu={float(self.nu):g}, forcing={self.forcing!r})"


def is_dealiased(u_hat, grid):
    """True when u_hat has no content outside the per-axis 2/3 box (the
    solver invariant); such states take the pruned transforms, which give
    the same result as the dense ones with about half the memory traffic."""
    u = check_spectral(u_hat, grid, 3)
    out = torch.empty(1, dtype=torch.float64, device=u.device)
    ws = grid.workspace(_lib.OP_REDUCE)
    _lib.call("hit3d_alias_energy", grid.desc, u.data_ptr(), out.data_ptr(), ws.data_ptr(),
              ws.numel(), stream())
    return float(out.item()) == 0.0


class NSRhs:
    """The callable returned by make_ns_rhs (navierstokes.py:62-97)."""

    def __init__(self, grid):
        self.grid = grid

    def __call__(self, u_hat, params, t, dealiased=None):
        """`dealiased` (bool) skips the per-call check (one host sync) when
        the caller knows whether u_hat is inside the 2/3 box."""
        g = self.grid
        u = check_spectral(u_hat, g, 3)
        if dealiased is None:
            dealiased = is_dealiased(u, g)
        pruned = 1 if dealiased else 0
        du = torch.empty_like(u)
        ws = g.workspace(_lib.OP_RHS, 3, pruned)
        f = params.forcing.desc()
        _lib.call("hit3d_rhs", g.desc, u.data_ptr(), du.data_ptr(), float(params.nu),
                  ctypes.byref(f), float(t), pruned, ws.data_ptr(), ws.numel(), stream())
        return du

    def __repr__(self):
        return f"ns_rhs({self.grid!r})"


def make_ns_rhs(grid):
    """navierstokes.py:62-97."""
    return NSRhs(grid)


def linear_operator(grid, params):
    """navierstokes.py:100-103: -nu k^2."""
    return -float(params.nu) * grid.k2
