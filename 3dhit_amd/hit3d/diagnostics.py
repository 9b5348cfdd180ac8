"""FP64 spectral diagnostics on the GPU (diagnostics.py of hit3d).

Every reduction runs in double precision on the device with a fixed
summation order (bit-reproducible), Neumaier-compensated across blocks.
Results come back to the host as Python floats / NumPy vectors, like the
reference's `energy_spectrum` (diagnostics.py:94-101).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .grid import DEALIASED_MARK, check_spectral, stream

__all__ = [
    "energy", "enstrophy", "dissipation", "energy_spectrum", "component_spectra",
    "energy_budget", "energy_budget_cb", "quadratic_sums", "shell_bins",
]


def quadratic_sums(u_hat, grid, nu=0.0, kmin=1.0, kmax=0.0):
    """Device FP64 sums -> host array [E, eps = 2 nu Omega, Omega, E_band]
    (diagnostics.py:56-70; forcing.py:48-52 for the band, kmin^2 <= k2 <= kmax^2)."""
    u = check_spectral(u_hat, grid, 3)
    out = torch.empty(4, dtype=torch.float64, device=u.device)
    ws = grid.workspace(_lib.OP_REDUCE)
    _lib.call("hit3d_energy_dissipation", grid.desc, u.data_ptr(), float(nu), float(kmin),
              float(kmax), out.data_ptr(), ws.data_ptr(), ws.numel(), stream())
    return out.cpu().numpy()


def energy(u_hat, grid):
    """diagnostics.py:56-58: E = sum w |u|^2 / (2 Ntot^2)."""
    return float(quadratic_sums(u_hat, grid)[0])


def enstrophy(u_hat, grid):
    """diagnostics.py:61-65: Omega = sum w k2 |u|^2 / (2 Ntot^2)."""
    return float(quadratic_sums(u_hat, grid)[2])


def dissipation(u_hat, grid, nu):
    """diagnostics.py:68-70: eps = 2 nu Omega."""
    return float(quadratic_sums(u_hat, grid, nu)[1])


def shell_bins(grid):
    n = ctypes.c_int64()
    _lib.call("hit3d_shell_bins", grid.desc, ctypes.byref(n))
    return int(n.value)


def shell_spectrum(u_hat, grid, component=-1):
    """All bins 0..smax (bin 0 = the k = 0 mode) as a device float64 tensor.
    A state the library marked as dealiased (see grid.mark_dealiased) is
    read only inside the 2/3 box; any other one in full (no device check)."""
    u = check_spectral(u_hat, grid, 3)
    nb = shell_bins(grid)
    out = torch.empty(nb, dtype=torch.float64, device=u.device)
    ws = grid.workspace(_lib.OP_SPECTRUM)
    box = 1 if getattr(u_hat, DEALIASED_MARK, None) == u_hat._version else 0
    _lib.call("hit3d_shell_spectrum", grid.desc, u.data_ptr(), int(component), box, out.data_ptr(),
              ws.data_ptr(), ws.numel(), stream())
    return out


def _ks(grid):
    k0 = 2 * np.pi / grid.Lx
    smax = shell_bins(grid) - 1
    return np.arange(1, smax + 1) * k0


def energy_spectrum(u_hat, grid):
    """diagnostics.py:94-101: (ks, E) with E[s-1] the energy of shell s."""
    return _ks(grid), shell_spectrum(u_hat, grid).cpu().numpy()[1:]


def component_spectra(u_hat, grid):
    """diagnostics.py:104-114: (ks, E11, E22, E33)."""
    return (_ks(grid), *[shell_spectrum(u_hat, grid, c).cpu().numpy()[1:] for c in range(3)])


def energy_budget(u_hat, grid, nu, forcing):
    """diagnostics.py:162-172: {E, eps, P}."""
    s = quadratic_sums(u_hat, grid, nu)
    return {"E": float(s[0]), "eps": float(s[1]), "P": float(forcing.injection(u_hat, grid))}


def energy_budget_cb(state):
    """diagnostics.py:175-182."""
    return energy_budget(state.u_hat, state.grid, state.params.nu, state.params.forcing)
