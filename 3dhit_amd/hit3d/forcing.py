"""Forcing plug-ins (forcing.py of hit3d) with the same contract:
``term(u_hat, grid, t)``, ``injection(u_hat, grid)``, ``label``.

Inside `rhs` / `rk4_step` / `rollout` the forcing never runs as separate
tensor code: its descriptor (`desc`) is handed to the fused HIP stage
kernel, and the coefficients that depend on the state (BandForcing's
eps/2E_band, LundgrenForcing's A) are reduced in FP64 on the device at every
RK4 stage. `term` and `injection` are the standalone API forms; `term` runs
the library's hit3d_forcing_term (coefficient and field on the device, no
host read).
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .diagnostics import quadratic_sums
from .grid import check_spectral, stream

__all__ = ["NoForcing", "LinearForcing", "BandForcing", "LundgrenForcing"]


def _fmt(x):
    return f"{float(x):g}"


def _tiny(grid):
    return float(np.finfo(np.float32 if grid.dtype == torch.float32 else np.float64).eps)


def _device_term(forcing, u_hat, grid):
    """f(u_hat) of `forcing` by hit3d_forcing_term (a new state tensor)."""
    u = check_spectral(u_hat, grid, 3)
    out = torch.empty_like(u)
    ws = grid.workspace(_lib.OP_REDUCE)
    f = forcing.desc()
    _lib.call("hit3d_forcing_term", grid.desc, u.data_ptr(), ctypes.byref(f), out.data_ptr(),
              ws.data_ptr(), ws.numel(), stream())
    return out


@dataclass(frozen=True)
class NoForcing:
    """forcing.py:55-71."""

    def desc(self):
        return _lib.ForcingDesc(_lib.FORCING_NONE, 0)

    def term(self, u_hat, grid, t):
        return torch.zeros((), dtype=u_hat.dtype, device=u_hat.device)

    def injection(self, u_hat, grid):
        return 0.0

    @property
    def label(self):
        return "NoForcing"

    def __repr__(self):
        return "NoForcing()"


@dataclass(frozen=True)
class LinearForcing:
    """forcing.py:74-106: f = A u, injection = 2 A E."""

    A: float

    def desc(self):
        return _lib.ForcingDesc(_lib.FORCING_LINEAR, 0, float(self.A))

    def term(self, u_hat, grid, t):
        return _device_term(self, u_hat, grid)

    def injection(self, u_hat, grid):
        return float(self.A) * 2.0 * float(quadratic_sums(u_hat, grid)[0])

    @property
    def label(self):
        return f"LinearForcing_A{_fmt(self.A)}"

    def __repr__(self):
        return f"LinearForcing(A={_fmt(self.A)})"


@dataclass(frozen=True)
class BandForcing:
    """forcing.py:109-163: f = eps / (2 E_band) u on kmin^2 <= k2 <= kmax^2,
    zero when 2 E_band <= finfo(dtype).eps."""

    eps: float
    kmin: float
    kmax: float

    def desc(self):
        return _lib.ForcingDesc(_lib.FORCING_BAND, 0, 0.0, float(self.eps), float(self.kmin),
                                float(self.kmax))

    def _two_e_band(self, u_hat, grid):
        return 2.0 * float(quadratic_sums(u_hat, grid, 0.0, self.kmin, self.kmax)[3])

    def term(self, u_hat, grid, t):
        return _device_term(self, u_hat, grid)

    def injection(self, u_hat, grid):
        return float(self.eps) if self._two_e_band(u_hat, grid) > _tiny(grid) else 0.0

    @property
    def label(self):
        return f"BandForcing_eps{_fmt(self.eps)}_k{_fmt(self.kmin)}-{_fmt(self.kmax)}"

    def __repr__(self):
        return f"BandForcing(eps={_fmt(self.eps)}, k=[{_fmt(self.kmin)}, {_fmt(self.kmax)}])"


@dataclass(frozen=True)
class LundgrenForcing:
    """Constant-energy linear forcing f = A(u) u with
    A(u) = (2 nu Omega(u) - (E(u) - energy)/tau) / (2 E(u)), zero when
    2E <= finfo(dtype).eps; tau=None keeps only the balance term, so that
    dE/dt = 2AE - eps = 0. Same contract as forcing.py:5-20; the FP64 energy
    and enstrophy sums that set A are taken on the stage state."""

    nu: float
    energy: float = 0.5
    tau: float | None = None

    def desc(self):
        tau = 0.0 if self.tau is None else float(self.tau)
        return _lib.ForcingDesc(_lib.FORCING_LUNDGREN, 0, 0.0, 0.0, 0.0, 0.0, float(self.nu),
                                float(self.energy), tau)

    def coefficient(self, u_hat, grid):
        s = quadratic_sums(u_hat, grid, self.nu)
        twoE, twoO = 2.0 * float(s[0]), 2.0 * float(s[2])
        if not twoE > _tiny(grid):
            return 0.0
        rate = float(self.nu) * twoO
        if self.tau is not None and self.tau > 0 and math.isfinite(self.tau):
            rate -= (0.5 * twoE - float(self.energy)) / float(self.tau)
        return rate / twoE

    def term(self, u_hat, grid, t):
        return _device_term(self, u_hat, grid)

    def injection(self, u_hat, grid):
        return self.coefficient(u_hat, grid) * 2.0 * float(quadratic_sums(u_hat, grid)[0])

    @property
    def label(self):
        tau = "inf" if self.tau is None else _fmt(self.tau)
        return f"LundgrenForcing_E{_fmt(self.energy)}_tau{tau}"

    def __repr__(self):
        return f"LundgrenForcing(nu={_fmt(self.nu)}, energy={_fmt(self.energy)}, tau={self.tau})"
