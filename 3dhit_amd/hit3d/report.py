"""Readers of recorded runs: the counterpart of `FieldWriter`'s sinks.

A run is read through a small view object over wherever the sink put it:

* an HDF5 file in the reference's layout (a path or an open h5py.File;
  needs h5py, imported on use), or
* a `MemorySink` (snapshots and series kept in host memory by
  `FieldWriter(sink=...)` and `save_series(sink, ...)`).

Every reader takes either, so the public functions of report.py:22-101 of
hit3d keep their names, results and messages (`load_run`, `read_series`,
`step_keys`, `read_grid`, `read_snapshot`) while no part of them depends on
the storage format.
"""

from __future__ import annotations

import contextlib
import warnings
from types import SimpleNamespace

import numpy as np
import torch

from .callbacks import MemorySink, _h5py
from .grid import make_grid
from .schema import grid_key, is_step_key, series_key, step_key, step_num

__all__ = ["load_run", "read_series", "step_keys", "read_grid", "read_snapshot"]

_GRID_FIELDS = ("Nx", "Ny", "Nz", "Lx", "Ly", "Lz")


class _FileView:
    """A recorded run inside an open HDF5 file (reference key layout)."""

    def __init__(self, f):
        self.f = f

    def steps(self):
        return sorted(step_num(k) for k in self.f.keys() if is_step_key(k))

    def grid_meta(self):
        if "grid" not in self.f:
            return None
        meta = {k: self.f[grid_key(k)][()] for k in _GRID_FIELDS}
        if grid_key("nu") in self.f:
            meta["nu"] = self.f[grid_key("nu")][()]
        return meta

    def field_info(self, step, field="u_hat"):
        d = self.f[f"{step_key(step)}/{field}"]
        return tuple(d.shape), np.dtype(d.dtype)

    def field(self, step, field="u_hat"):
        return np.asarray(self.f[f"{step_key(step)}/{field}"])

    def series(self, name):
        grp = series_key(name)
        if grp not in self.f:
            return None
        return {k: np.asarray(self.f[grp][k]) for k in self.f[grp].keys()}


class _SinkView:
    """A recorded run held by a MemorySink."""

    def __init__(self, sink):
        self.sink = sink

    def steps(self):
        return sorted(self.sink.snapshots)

    def grid_meta(self):
        return None if self.sink.meta is None else dict(self.sink.meta)

    def field_info(self, step, field="u_hat"):
        a = self.sink.snapshots[step][field]
        return tuple(a.shape), np.dtype(a.dtype)

    def field(self, step, field="u_hat"):
        return np.asarray(self.sink.snapshots[step][field])

    def series(self, name):
        s = self.sink.series.get(name)
        return None if s is None else {k: np.asarray(v) for k, v in s.items()}


def _is_h5_file(obj):
    return type(obj).__module__.startswith("h5py") and hasattr(obj, "keys")


@contextlib.contextmanager
def _view(source):
    """View over `source`: a MemorySink, an open h5py.File or a file path."""
    if isinstance(source, MemorySink):
        yield _SinkView(source)
    elif _is_h5_file(source):
        yield _FileView(source)
    else:
        with _h5py().File(source, "r") as f:
            yield _FileView(f)


def _describe(source):
    return "the memory sink" if isinstance(source, MemorySink) else str(getattr(source, "filename", source))


def _grid_of(view, dtype):
    meta = view.grid_meta()
    if meta is not None:
        g = make_grid(int(meta["Nx"]), int(meta["Ny"]), int(meta["Nz"]), Lx=float(meta["Lx"]),
                      Ly=float(meta["Ly"]), Lz=float(meta["Lz"]), dtype=dtype)
        return g, (float(meta["nu"]) if "nu" in meta else None)
    steps = view.steps()
    if not steps:
        raise ValueError("no snapshots found in file")
    _, nx, ny, nzh = view.field_info(steps[0])[0]
    nz = 2 * (nzh - 1)                       # the rfft layout of an even Nz
    warnings.warn(f"no grid metadata in file; assuming a 2π cube ({nx}x{ny}x{nz}) and unknown nu",
                  stacklevel=3)
    return make_grid(nx, ny, nz, dtype=dtype), None


def step_keys(f):
    """Sorted step_######## group names of a run (open file or MemorySink)."""
    with _view(f) as v:
        return [step_key(n) for n in v.steps()]


def read_grid(f, dtype):
    """(Grid of precision `dtype`, nu) from the run's grid metadata. A run
    recorded without it is read as a 2 pi cube with unknown nu whose sizes
    come from the first snapshot's rfft layout (even Nz), with a warning."""
    with _view(f) as v:
        return _grid_of(v, dtype)


def read_snapshot(f, step, field="u_hat"):
    """One recorded field (default the spectral state) of snapshot `step`, as NumPy."""
    with _view(f) as v:
        return v.field(int(step), field)


def load_run(source):
    """Namespace(grid, nu, dtype, steps) describing a recorded run without
    reading any field array; ValueError when it holds no snapshot."""
    with _view(source) as v:
        steps = v.steps()
        if not steps:
            raise ValueError(f"no snapshots found in {_describe(source)}")
        _, cdtype = v.field_info(steps[0])
        dtype = torch.float64 if cdtype == np.complex128 else torch.float32
        grid, nu = _grid_of(v, dtype)
        return SimpleNamespace(grid=grid, nu=nu, dtype=dtype, steps=steps)


def read_series(source, name):
    """The series recorded under `name` (save_series) as a dict of NumPy
    arrays, `t` plus `values` or one vector per recorded key; None when the
    run has no series of that name."""
    with _view(source) as v:
        return v.series(name)
