"""Shared test setup: import paths for the oracle and the `hit3d` package,
the `gpu` marker, and small field builders used across modules."""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "3dhit_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


@pytest.fixture
def grid32():
    import hit3d_oracle as O
    return O.make_grid(32, dtype=np.float64)


def fill3(f, grid):
    """4D physical field with the same 3D pattern in every component
    (reference tests/conftest.py:38-40)."""
    return np.broadcast_to(f, (3, grid.Nx, grid.Ny, grid.Nz))


def dealiased_random_spectral(seed, grid):
    """rfft of a random real field, dealiased (reference tests/conftest.py:43-48)."""
    import hit3d_oracle as O
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((3, grid.Nx, grid.Ny, grid.Nz)).astype(grid.dtype)
    return O.dealias(O.to_spectral(u, grid), grid)
