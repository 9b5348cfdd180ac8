"""Time integration (integrators.py of hit3d).

`rollout` runs all `nsteps` classic RK4 steps inside the library
(`hit3d_rk4_steps`): per stage five pencil-pass kernels, with the stage
combination and the FP64 forcing reductions fused into the last pass; the
host only enqueues. `evolve` is the host-side observation driver over
`rollout` segments.
"""

from __future__ import annotations

import ctypes
import math
import time

import torch

from . import _lib
from .grid import check_spectral, stream
from .navierstokes import is_dealiased

__all__ = ["rk4_step", "rollout", "evolve", "State", "Callback", "Diagnostic"]


def _steps(rhs, u_hat, params, dt, nsteps, t0, diag=None):
    g = rhs.grid
    u = check_spectral(u_hat, g, 3).clone()
    if nsteps == 0:
        return u
    pruned = 1 if is_dealiased(u, g) else 0
    ws = g.workspace(_lib.OP_RK4, 3, pruned)
    f = params.forcing.desc()
    _lib.call("hit3d_rk4_steps", g.desc, u.data_ptr(), float(params.nu), float(dt), int(nsteps),
              float(t0), ctypes.byref(f), pruned,
              None if diag is None else diag.data_ptr(), ws.data_ptr(), ws.numel(), stream())
    return u


def rk4_step(rhs, u_hat, params, dt, t):
    """integrators.py:32-38: one classic RK4 step."""
    return _steps(rhs, u_hat, params, dt, 1, t)


def rollout(rhs, u_hat, params, dt, nsteps, t0=0.0):
    """integrators.py:41-53: `nsteps` RK4 steps from t0 (pure: the input is
    not modified)."""
    if int(nsteps) < 0:
        raise ValueError("nsteps must be >= 0")
    return _steps(rhs, u_hat, params, dt, int(nsteps), t0)


# ---------------------------------------------------------------------------
# Observation driver (integrators.py:56-109, callbacks.py:33-105)
# ---------------------------------------------------------------------------


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
    """callbacks.py:69-105: records float(t) and f(state) every `every`
    steps (values pulled to the host). HDF5 persistence (path/name) is not
    part of this package."""

    def __init__(self, f, *, every=1, path=None, name=None):
        if path is not None and name is None:
            raise ValueError("Diagnostic(..., path=...) also needs `name`")
        if path is not None:
            raise NotImplementedError("HDF5 series persistence is not implemented")
        self.f, self.every, self.path, self.name = f, every, path, name
        self.times, self.values = [], []

    def is_due(self, step, t, dt):
        return step % self.every == 0

    def fire(self, state):
        self.times.append(float(state.t))
        self.values.append(self.f(state))

    def finalize(self):
        pass


def evolve(rhs, u_hat, params, grid, dt, nsteps, *, t0=0.0, callbacks=(), progress=False):
    """integrators.py:56-109: host loop over rollout segments, firing the
    callbacks that are due at segment boundaries."""
    if isinstance(progress, bool):
        print_every = max(1, nsteps // 20) if progress else 0
    else:
        print_every = max(int(progress), 1)

    def due_any(n):
        return any(cb.is_due(n, t0 + n * dt, dt) for cb in callbacks)

    stops = sorted({n for n in range(1, nsteps + 1) if due_any(n)}
                   | (set(range(print_every, nsteps + 1, print_every)) if print_every else set())
                   | {nsteps}) if nsteps > 0 else []

    def fire(n, u):
        t = t0 + n * dt
        state = State(u, t, n, params, grid)
        for cb in callbacks:
            if cb.is_due(n, t, dt):
                cb.fire(state)

    wall0 = time.time()
    fire(0, u_hat)
    prev = 0
    for n in stops:
        u_hat = rollout(rhs, u_hat, params, dt, n - prev, t0 + prev * dt)
        fire(n, u_hat)
        if print_every and (n % print_every == 0 or n == nsteps):
            el = time.time() - wall0
            print(f"step {n}/{nsteps}  t = {t0 + n * dt:<9.4g}  elapsed {el:.1f}s", flush=True)
        prev = n
    for cb in callbacks:
        cb.finalize()
    return u_hat
