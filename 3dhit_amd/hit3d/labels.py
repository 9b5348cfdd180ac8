"""Filesystem-safe run labels (labels.py:1-31 of hit3d): `label(x)` of a
grid, forcing or NSParams, and `run_label(grid, params, scheme)` =
"<grid>_<params>_<scheme>". Numbers print with %g, so float32 and float64
parameters give the same label."""

from __future__ import annotations

from .grid import Grid

__all__ = ["label", "run_label"]


def label(x):
    if isinstance(x, Grid):
        if x.Nx == x.Ny == x.Nz:
            return f"N{x.Nx}"
        return f"N{x.Nx}x{x.Ny}x{x.Nz}"
    return x.label


def run_label(grid, params, scheme="RK4"):
    return f"{label(grid)}_{label(params)}_{scheme}"
