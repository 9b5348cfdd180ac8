"""Readers of `FieldWriter` snapshot files (report.py:22-101 of hit3d):
`load_run`, `read_series`, `step_keys`, `read_grid`, `read_snapshot`.
They need h5py (imported when called; there is no other file format)."""

from __future__ import annotations

import warnings
from types import SimpleNamespace

import numpy as np
import torch

from .callbacks import _h5py
from .grid import make_grid
from .schema import grid_key, is_step_key, series_key, step_key, step_num

__all__ = ["load_run", "read_series", "step_keys", "read_grid", "read_snapshot"]


def step_keys(f):
    """Sorted step_######## group names of an open file."""
    return sorted(k for k in f.keys() if is_step_key(k))


def read_grid(f, dtype):
    """The Grid (precision `dtype`) and nu from the file's grid/* group; a
    file without it is taken as a 2 pi cube with unknown nu, sized from the
    rfft layout of its first snapshot (even Nz assumed), with a warning."""
    if "grid" in f:
        g = make_grid(int(f[grid_key("Nx")][()]), int(f[grid_key("Ny")][()]),
                      int(f[grid_key("Nz")][()]), Lx=float(f[grid_key("Lx")][()]),
                      Ly=float(f[grid_key("Ly")][()]), Lz=float(f[grid_key("Lz")][()]),
                      dtype=dtype)
        nu = float(f[grid_key("nu")][()]) if grid_key("nu") in f else None
        return g, nu
    keys = step_keys(f)
    if not keys:
        raise ValueError("no snapshots found in file")
    shape = f[f"{keys[0]}/u_hat"].shape
    nz = 2 * (shape[3] - 1)
    warnings.warn(f"no grid metadata in file; assuming a 2π cube ({shape[1]}x{shape[2]}x{nz}) "
                  "and unknown nu", stacklevel=2)
    return make_grid(shape[1], shape[2], nz, dtype=dtype), None


def _state_dtype(f):
    c = f[f"{step_keys(f)[0]}/u_hat"].dtype
    return torch.float64 if c == np.complex128 else torch.float32


def read_snapshot(f, step):
    """Spectral state of snapshot `step` (NumPy)."""
    return np.asarray(f[f"{step_key(step)}/u_hat"])


def load_run(path):
    """Namespace(grid, nu, dtype, steps) of the file at `path` (no field
    arrays are read); errors if it holds no snapshot."""
    h5py = _h5py()
    with h5py.File(path, "r") as f:
        keys = step_keys(f)
        if not keys:
            raise ValueError(f"no snapshots found in {path}")
        dtype = _state_dtype(f)
        grid, nu = read_grid(f, dtype)
        return SimpleNamespace(grid=grid, nu=nu, dtype=dtype, steps=[step_num(k) for k in keys])


def read_series(path_or_file, name):
    """series/<name> as a dict of NumPy arrays (t plus `values` or one
    vector per key), or None if the file has no such group."""
    h5py = _h5py()
    if isinstance(path_or_file, h5py.File):
        return _read_series(path_or_file, name)
    with h5py.File(path_or_file, "r") as f:
        return _read_series(f, name)


def _read_series(f, name):
    grp = series_key(name)
    if grp not in f:
        return None
    return {k: np.asarray(f[grp][k]) for k in f[grp].keys()}
