"""Observation layer of `evolve` (callbacks.py:1-217 of hit3d).

Callback functions are pure, f(state) -> value, on the read-only `State`
(u_hat, t, step, params, grid); scheduling (`every` steps or `every_time`)
and all accumulation / IO belong to the wrappers. Everything fires at step 0
as well. Nothing here touches `rollout`, the differentiable path.

Snapshots (`FieldWriter`) leave the device asynchronously: the copy of each
field runs on a side HIP stream into a page-locked host buffer that is
reused from snapshot to snapshot, ordered after the producing stream by an
event, and a writer thread hands the host copy to the sink (HDF5 by default)
once the copy's event has completed. The solver's stream never waits on
the PCIe transfer or the file system; at most `depth` snapshots are in
flight (the next `fire` blocks until a slot is free), which bounds the
page-locked memory to (depth + 1) copies of the written fields.
"""

from __future__ import annotations

import math
import os
import queue
import threading

import numpy as np
import torch

from .schema import grid_key, series_field, series_key, step_key

__all__ = ["State", "Callback", "Diagnostic", "FieldWriter", "save_series", "Hdf5Sink",
           "MemorySink"]


def _h5py():
    try:
        import h5py
    except ImportError as e:  # the HDF5 layer has no fallback format
        raise ImportError("HDF5 snapshot files need h5py, which is not installed; "
                          "pass a `sink` to FieldWriter to keep snapshots another way") from e
    return h5py


class State(tuple):
    """callbacks.py:33-40: read-only (u_hat, t, step, params, grid)."""

    __slots__ = ()
    _fields = ("u_hat", "t", "step", "params", "grid")

    def __new__(cls, u_hat, t, step, params, grid):
        return tuple.__new__(cls, (u_hat, t, step, params, grid))

    u_hat = property(lambda s: s[0])
    t = property(lambda s: s[1])
    step = property(lambda s: s[2])
    params = property(lambda s: s[3])
    grid = property(lambda s: s[4])


class Callback:
    """callbacks.py:43-66: calls f(state) every `every` steps, or (when
    `every_time` is set) at step 0 and whenever t crosses a multiple of
    `every_time`."""

    def __init__(self, f, *, every=1, every_time=None):
        self.f, self.every, self.every_time = f, every, every_time

    def is_due(self, step, t, dt):
        if self.every_time is None:
            return step % self.every == 0
        return step == 0 or math.floor(t / self.every_time) > math.floor((t - dt) / self.every_time)

    def fire(self, state):
        self.f(state)

    def finalize(self):
        pass


class Diagnostic:
    """callbacks.py:69-105: records float(t) and f(state) (a scalar or a dict
    of scalars, pulled to the host) every `every` steps. With `path` and
    `name`, `finalize` (called by `evolve` at the end of the run) stores the
    series under series/<name> of the HDF5 file at `path`."""

    def __init__(self, f, *, every=1, path=None, name=None):
        if path is not None and name is None:
            raise ValueError("Diagnostic(..., path=...) also needs `name` "
                             "(the series/<name> group to persist it under)")
        if path is not None and not isinstance(path, MemorySink):
            _h5py()
        self.f, self.every, self.path, self.name = f, every, path, name
        self.times, self.values = [], []

    def is_due(self, step, t, dt):
        return step % self.every == 0

    def fire(self, state):
        self.times.append(float(state.t))
        v = self.f(state)
        if isinstance(v, torch.Tensor):
            v = v.item() if v.numel() == 1 else v.detach().cpu().numpy()
        elif isinstance(v, dict):
            v = {k: (x.item() if isinstance(x, torch.Tensor) else x) for k, x in v.items()}
        self.values.append(v)

    def finalize(self):
        if self.path is not None:
            save_series(self.path, self.name, self)


def _series_columns(diagnostic):
    """t plus `values` (scalar diagnostics) or one column per key (dict values)."""
    cols = {"t": np.asarray(diagnostic.times)}
    values = diagnostic.values
    if values and isinstance(values[0], dict):
        for fld in values[0]:
            cols[fld] = np.asarray([v[fld] for v in values])
    else:
        cols["values"] = np.asarray(values)
    return cols


def save_series(path, name, diagnostic):
    """callbacks.py:188-217: append a Diagnostic's series as series/<name>/t
    plus `values` (scalars) or one vector per key (dict values), to the HDF5
    file at `path` or to a MemorySink. Errors if series/<name> already exists."""
    cols = _series_columns(diagnostic)
    if isinstance(path, MemorySink):
        if name in path.series:
            raise ValueError(f"{series_key(name)} already exists in the memory sink (saving the "
                             "same series twice? reruns should start from a fresh sink)")
        path.series[name] = cols
        return path
    h5py = _h5py()
    with h5py.File(path, "a") as f:
        if series_key(name) in f:
            raise ValueError(f"{series_key(name)} already exists in {path} (saving the same "
                             "series twice? reruns should start from a fresh file)")
        for fld, col in cols.items():
            f[series_field(name, fld)] = col
    return path


def _grid_meta(state):
    g = state.grid
    meta = {"Nx": g.Nx, "Ny": g.Ny, "Nz": g.Nz, "Lx": g.Lx, "Ly": g.Ly, "Lz": g.Lz}
    nu = getattr(state.params, "nu", None)
    if nu is not None:
        meta["nu"] = float(nu)
    return meta


class Hdf5Sink:
    """Writes snapshots in the reference file layout: grid/* on the first
    write, then step_<n>/t and step_<n>/<field> (callbacks.py:169-182)."""

    def __init__(self, path):
        self.h5py = _h5py()
        self.path = path

    def write(self, step, t, arrays, meta):
        with self.h5py.File(self.path, "a") as f:
            if "grid" not in f:
                for k, v in meta.items():
                    f[grid_key(k)] = v
            grp = step_key(step)
            f[f"{grp}/t"] = float(t)
            for name, arr in arrays.items():
                f[f"{grp}/{name}"] = arr


class MemorySink:
    """Keeps a run in host memory: `snapshots[step] = {"t": t, field:
    array}`, the grid metadata in `meta` and the series that save_series
    (or Diagnostic(path=sink, name=...)) stores in `series[name]`. The
    report readers (load_run, read_series, ...) accept it in place of a file."""

    def __init__(self):
        self.snapshots, self.meta, self.series = {}, None, {}

    def write(self, step, t, arrays, meta):
        if self.meta is None:
            self.meta = dict(meta)
        self.snapshots[int(step)] = {"t": float(t), **{k: np.array(v) for k, v in arrays.items()}}


class _AsyncCopier:
    """Device -> page-locked host copies on a side stream, drained by one
    writer thread into a sink."""

    def __init__(self, sink, depth=2):
        self.sink = sink
        self.q = queue.Queue(maxsize=max(int(depth), 1))
        self.pool = {}
        self.lock = threading.Lock()
        self.err = None
        self.stream = None
        self.thread = None

    def _take(self, t):
        key = (tuple(t.shape), t.dtype)
        with self.lock:
            free = self.pool.setdefault(key, [])
            if free:
                return free.pop()
        return torch.empty(t.shape, dtype=t.dtype, pin_memory=True)

    def _give(self, buf):
        with self.lock:
            self.pool.setdefault((tuple(buf.shape), buf.dtype), []).append(buf)

    def submit(self, step, t, arrays, meta):
        if self.err is not None:
            raise self.err
        if self.thread is None:
            self.thread = threading.Thread(target=self._run, name="hit3d-snapshots", daemon=True)
            self.thread.start()
        host, done = {}, None
        for name, val in arrays.items():
            if isinstance(val, torch.Tensor) and val.is_cuda:
                if self.stream is None:
                    self.stream = torch.cuda.Stream(device=val.device)  # default priority
                ready = torch.cuda.Event()
                ready.record(torch.cuda.current_stream(val.device))
                self.stream.wait_event(ready)
                buf = self._take(val)
                with torch.cuda.stream(self.stream):
                    buf.copy_(val, non_blocking=True)
                val.record_stream(self.stream)  # keep the allocator off it until copied
                host[name] = buf
            elif isinstance(val, torch.Tensor):
                host[name] = val.detach().numpy().copy()
            else:
                host[name] = np.asarray(val)
        if self.stream is not None:
            done = torch.cuda.Event()
            done.record(self.stream)
        self.q.put((step, t, host, dict(meta), done))

    def _run(self):
        while True:
            item = self.q.get()
            if item is None:
                self.q.task_done()
                return
            step, t, host, meta, done = item
            try:
                if done is not None:
                    done.synchronize()
                arrays = {k: (v.numpy() if isinstance(v, torch.Tensor) else v) for k, v in host.items()}
                self.sink.write(step, t, arrays, meta)
            except BaseException as e:  # reported by the next submit / flush
                self.err = e
            finally:
                for v in host.values():
                    if isinstance(v, torch.Tensor):
                        self._give(v)
                self.q.task_done()

    def flush(self):
        if self.thread is not None:
            self.q.join()
        if self.err is not None:
            err, self.err = self.err, None
            raise err

    def close(self):
        if self.thread is not None:
            self.q.put(None)
            self.thread.join()
            self.thread = None


class FieldWriter:
    """callbacks.py:108-185: full-field snapshots of the state fields in
    `fields` (with the time) under step_<n>/<field>, plus a grid/* group
    (Nx..Lz and the params' nu) on the first write, in an HDF5 file at
    `path`. `overwrite=True` deletes an existing file at construction.

    Scheduled by step count (`every`) or, with `every_time`, by simulation
    time: first at t = warmup_time (default 0), then every `every_time`
    (nothing before warmup_time).

    Writes are asynchronous (device -> pinned host on a side stream, then a
    writer thread); `finalize` (called by `evolve`) waits for all of them
    and re-raises a failed write. `sink` replaces the HDF5 file by any
    object with write(step, t, arrays, meta) (e.g. `MemorySink`).
    """

    def __init__(self, path=None, *, every=1, fields=("u_hat",), overwrite=True,
                 every_time=None, warmup_time=None, sink=None, depth=2):
        if every_time is not None and warmup_time is None:
            warmup_time = 0.0
        if sink is None:
            if path is None:
                raise ValueError("FieldWriter needs a `path` (HDF5 file) or a `sink`")
            sink = Hdf5Sink(path)
        if path is not None:
            d = os.path.dirname(path)
            if d:
                os.makedirs(d, exist_ok=True)
            if overwrite and os.path.exists(path):
                os.remove(path)
        self.path, self.every, self.fields = path, every, tuple(fields)
        self.every_time, self.warmup_time = every_time, warmup_time
        self.sink = sink
        self._copier = _AsyncCopier(sink, depth)

    def is_due(self, step, t, dt):
        if self.every_time is None:
            return step % self.every == 0
        if t < self.warmup_time:
            return False
        # cadence indices within dt/(4 every_time) of an integer are snapped
        # to it, so float noise in t cannot count one crossing twice
        tol = (dt / self.every_time) / 4

        def index(x):
            r = round(x)
            return r if abs(x - r) < tol else x

        n_now = math.floor(index((t - self.warmup_time) / self.every_time))
        n_prev = math.floor(index((t - dt - self.warmup_time) / self.every_time))
        return n_now > n_prev

    def fire(self, state):
        arrays = {fld: getattr(state, fld) for fld in self.fields}
        self._copier.submit(state.step, state.t, arrays, _grid_meta(state))

    def finalize(self):
        try:
            self._copier.flush()
        finally:
            self._copier.close()
