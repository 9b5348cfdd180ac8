"""HDF5 key naming of the snapshot files that `FieldWriter` / `save_series`
write and `report` reads (schema.py:26-57 of hit3d; the on-disk layout is
the reference's, so files are interchangeable):

    grid/Nx, grid/Ny, grid/Nz, grid/Lx, grid/Ly, grid/Lz, grid/nu (optional)
    step_00000000/t, step_00000000/<field>, ...   one group per snapshot
    series/<name>/t, series/<name>/<field>, ...   one group per Diagnostic
"""

from __future__ import annotations

__all__ = ["STEP_PREFIX", "STEP_DIGITS", "step_key", "step_num", "is_step_key", "grid_key",
           "series_key", "series_field"]

STEP_PREFIX = "step_"
STEP_DIGITS = 8


def step_key(n):
    """Snapshot group of step n: step_key(50) == "step_00000050"."""
    return f"{STEP_PREFIX}{int(n):0{STEP_DIGITS}d}"


def is_step_key(key):
    return key.startswith(STEP_PREFIX)


def step_num(key):
    """Inverse of step_key."""
    return int(key.rsplit("_", 1)[-1])


def grid_key(name):
    """Grid metadata entry ("Nx", "Lx", "nu", ...)."""
    return f"grid/{name}"


def series_key(name):
    """Group of the dense time series recorded under `name`."""
    return f"series/{name}"


def series_field(name, fld):
    return f"{series_key(name)}/{fld}"
