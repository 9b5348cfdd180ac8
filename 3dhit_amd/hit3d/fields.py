"""Initial conditions (fields.py of hit3d), generated on the GPU.

`random_field` follows fields.py:35-68 step by step: uniform noise in
physical space -> rfft -> dealias -> project -> per-shell rescale. The noise
comes from Philox4x64-10 keyed by an integer seed (the reference's threefry
stream is not reproducible outside JAX); every step is a HIP kernel, and only
the vector of smax+1 shell factors is formed on the host.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .diagnostics import shell_spectrum
from .grid import check_spectral, mark_dealiased, stream, to_spectral

__all__ = ["spectral_state", "random_field", "taylor_green", "default_spectrum",
           "config_spectrum", "uniform_noise", "CONFIG_SPECTRUM_C"]

# C k^4 exp(-2 (k/4)^2) carries E = 0.5 on every N >= 64 grid with this C.
CONFIG_SPECTRUM_C = 0.004155648754181591


def spectral_state(grid):
    """fields.py:25-27: zero state (3, Nx, Ny, Nz//2+1)."""
    return torch.zeros((3, grid.Nx, grid.Ny, grid.Nz // 2 + 1), dtype=grid.complex_dtype,
                       device=grid.device)


def default_spectrum(k):
    """fields.py:30-32: k^4 exp(-2 k^2)."""
    return k**4 * np.exp(-2 * k**2)


def config_spectrum(k):
    """C k^4 exp(-2 (k/4)^2), total energy 0.5."""
    return CONFIG_SPECTRUM_C * k**4 * np.exp(-2 * (k / 4.0) ** 2)


def uniform_noise(seed, grid):
    """U[-0.5, 0.5) noise (3, Nx, Ny, Nz) in the grid dtype: value i of the
    flattened array is word i % 4 of Philox4x64-10 at counter i // 4 + 1,
    key (seed, 0)."""
    u = torch.empty((3, grid.Nx, grid.Ny, grid.Nz), dtype=grid.dtype, device=grid.device)
    _lib.call("hit3d_uniform_noise", grid.desc, int(seed) & (2**64 - 1), u.data_ptr(), stream())
    return u


def random_field(seed, grid, spectrum=default_spectrum):
    """fields.py:35-68 with a seed instead of a jax key. `spectrum` is a
    NumPy-compatible callable of the shell wavenumbers."""
    u_hat = to_spectral(uniform_noise(seed, grid), grid)
    _lib.call("hit3d_dealias_project", grid.desc, u_hat.data_ptr(), u_hat.data_ptr(), 3, 1, 1,
              stream())
    mark_dealiased(u_hat)                                    # the shell rescale keeps it so
    Es = shell_spectrum(u_hat, grid).cpu().numpy()           # bins 0..smax, FP64
    npd = np.float32 if grid.dtype == torch.float32 else np.float64
    k0 = 2 * np.pi / grid.Lx
    ks = np.arange(len(Es), dtype=npd) * npd(k0)
    target = np.asarray(spectrum(ks), dtype=np.float64)
    safe = np.where(Es > 0, Es, 1.0)
    c = np.where(Es > 0, np.where(target > 0, np.sqrt(target / safe), 0.0), 1.0).astype(npd)
    c_dev = torch.from_numpy(c.astype(np.float64)).to(grid.device)
    _lib.call("hit3d_shell_scale", grid.desc, u_hat.data_ptr(), 3, c_dev.data_ptr(), len(c),
              stream())
    return mark_dealiased(u_hat)


def taylor_green(grid):
    """fields.py:71-87 (the physical field is built with torch, then
    transformed and dealiased by the HIP kernels)."""
    t, d = grid.dtype, grid.device
    x = (torch.arange(grid.Nx, dtype=t, device=d) * (2 * np.pi / grid.Nx)).reshape(-1, 1, 1)
    y = (torch.arange(grid.Ny, dtype=t, device=d) * (2 * np.pi / grid.Ny)).reshape(1, -1, 1)
    z = (torch.arange(grid.Nz, dtype=t, device=d) * (2 * np.pi / grid.Nz)).reshape(1, 1, -1)
    u = torch.stack((torch.sin(x) * torch.cos(y) * torch.cos(z),
                     -torch.cos(x) * torch.sin(y) * torch.cos(z),
                     torch.zeros((grid.Nx, grid.Ny, grid.Nz), dtype=t, device=d)))
    u_hat = to_spectral(u.contiguous(), grid)
    _lib.call("hit3d_dealias_project", grid.desc, u_hat.data_ptr(), u_hat.data_ptr(), 3, 1, 0,
              stream())
    return mark_dealiased(u_hat)
