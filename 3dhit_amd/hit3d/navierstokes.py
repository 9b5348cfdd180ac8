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
from .grid import DEALIASED_MARK, check_spectral, mark_dealiased, stream

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
    the same result as the dense ones with about half the memory traffic.
    One FP64 reduction on the device and one host read of its result."""
    u = check_spectral(u_hat, grid, 3)
    out = torch.empty(1, dtype=torch.float64, device=u.device)
    ws = grid.workspace(_lib.OP_REDUCE)
    _lib.call("hit3d_alias_energy", grid.desc, u.data_ptr(), out.data_ptr(), ws.data_ptr(),
              ws.numel(), stream())
    return float(out.item()) == 0.0


def known_dealiased(u_hat, grid):
    """is_dealiased without the host read when the tensor carries a valid
    mark (same object, not written in place since it was marked); otherwise
    the device check, whose positive answer is then recorded."""
    if getattr(u_hat, DEALIASED_MARK, None) == u_hat._version:
        return True
    if is_dealiased(u_hat, grid):
        mark_dealiased(u_hat)
        return True
    return False


class NSRhs:
    """The callable returned by make_ns_rhs (navierstokes.py:62-97)."""

    def __init__(self, grid):
        self.grid = grid

    def __call__(self, u_hat, params, t, dealiased=None):
        """`dealiased` (bool) states whether u_hat is inside the 2/3 box; by
        default a marked tensor (see `mark_dealiased`) is taken as such and
        any other one is checked on the device (one host read)."""
        g = self.grid
        u = check_spectral(u_hat, g, 3)
        if dealiased is None:
            dealiased = known_dealiased(u_hat, g)
        pruned = 1 if dealiased else 0
        du = torch.empty_like(u)
        ws = g.workspace(_lib.OP_RHS, 3, pruned)
        f = params.forcing.desc()
        _lib.call("hit3d_rhs", g.desc, u.data_ptr(), du.data_ptr(), float(params.nu),
                  ctypes.byref(f), float(t), pruned, ws.data_ptr(), ws.numel(), stream())
        return mark_dealiased(du) if pruned else du

    def __repr__(self):
        return f"ns_rhs({self.grid!r})"


def make_ns_rhs(grid):
    """navierstokes.py:62-97."""
    return NSRhs(grid)


def linear_operator(grid, params):
    """navierstokes.py:100-103: -nu k^2 (the diagonal viscous operator that
    the integrating-factor scheme, rollout(..., scheme="if-rk4"), integrates
    exactly; its kernels form exp(-nu k2 dt) per mode from the same k2)."""
    return -float(params.nu) * grid.k2
