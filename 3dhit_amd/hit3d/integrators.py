"""Time integration (integrators.py of hit3d).

`rollout` runs all `nsteps` classic RK4 steps inside the library
(`hit3d_rk4_steps`): per stage five pencil-pass kernels, with the stage
combination and the FP64 forcing reductions fused into the last pass; the
host only enqueues. `evolve` is the host-side observation driver over
`rollout` segments.
"""

from __future__ import annotations

import ctypes
import time

from . import _lib
from .callbacks import Callback, Diagnostic, State
from .grid import check_spectral, mark_dealiased, stream
from .navierstokes import known_dealiased

__all__ = ["rk4_step", "if_rk4_step", "rollout", "evolve", "State", "Callback", "Diagnostic"]


_SCHEMES = {"rk4": "hit3d_rk4_steps", "if-rk4": "hit3d_rk4_if_steps"}


def _steps(rhs, u_hat, params, dt, nsteps, t0, diag=None, scheme="rk4"):
    """`nsteps` steps of hit3d_rk4_steps on a copy of u_hat. The layout
    (pruned when u_hat has no content outside the 2/3 box) is decided from
    the tensor's mark when it carries one, so a chain of rollouts on the
    library's own outputs never waits for the device."""
    if scheme not in _SCHEMES:
        raise ValueError(f"scheme must be one of {sorted(_SCHEMES)}, got {scheme!r}")
    g = rhs.grid
    inside = known_dealiased(u_hat, g)
    u = check_spectral(u_hat, g, 3).clone()
    if nsteps == 0:
        return mark_dealiased(u) if inside else u
    pruned = 1 if inside else 0
    ws = g.workspace(_lib.OP_RK4, 3, pruned)
    f = params.forcing.desc()
    _lib.call(_SCHEMES[scheme], g.desc, u.data_ptr(), float(params.nu), float(dt), int(nsteps),
              float(t0), ctypes.byref(f), pruned,
              None if diag is None else diag.data_ptr(), ws.data_ptr(), ws.numel(), stream())
    return mark_dealiased(u) if pruned else u


def rk4_step(rhs, u_hat, params, dt, t):
    """integrators.py:32-38: one classic RK4 step."""
    return _steps(rhs, u_hat, params, dt, 1, t)


def rollout(rhs, u_hat, params, dt, nsteps, t0=0.0, *, scheme="rk4"):
    """integrators.py:41-53: `nsteps` RK4 steps from t0 (pure: the input is
    not modified). scheme="if-rk4" takes the integrating-factor (Lawson) form
    instead: the viscous operator linear_operator(grid, params) = -nu k2 is
    integrated exactly and classic RK4 advances the rest (the option the
    reference leaves open at integrators.py:14-16)."""
    if int(nsteps) < 0:
        raise ValueError("nsteps must be >= 0")
    return _steps(rhs, u_hat, params, dt, int(nsteps), t0, scheme=scheme)


def if_rk4_step(rhs, u_hat, params, dt, t):
    """One integrating-factor (Lawson) RK4 step: with E_h = exp(-nu k2 dt/2)
    and N = rhs + nu k2 (the non-viscous part of the RHS),
    u2 = E_h (u + dt/2 k1), u3 = E_h u + dt/2 k2, u4 = E_h^2 u + dt E_h k3,
    u_new = E_h^2 u + dt/6 (E_h^2 k1 + 2 E_h (k2 + k3) + k4), k_i = N(u_i)."""
    return _steps(rhs, u_hat, params, dt, 1, t, scheme="if-rk4")


# ---------------------------------------------------------------------------
# Observation driver (integrators.py:56-123)
# ---------------------------------------------------------------------------


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
            _print_progress(n, nsteps, t0 + n * dt, wall0)
        prev = n
    for cb in callbacks:
        cb.finalize()
    return u_hat


def _print_progress(n, nsteps, t, wall0):
    """Progress line of integrators.py:112-120: step, percentage, simulation
    time, wall time and the linear ETA."""
    elapsed = time.time() - wall0
    eta = elapsed / n * (nsteps - n)
    width = len(str(nsteps))
    print(f"step {n:{width}d}/{nsteps} ({round(100 * n / nsteps):3d}%)  "
          f"t = {t:<9.4g}  elapsed {_fmt_secs(elapsed)}  eta {_fmt_secs(eta)}", flush=True)


def _fmt_secs(s):
    if s < 60:
        return f"{s:.1f}s"
    return f"{int(s // 60)}m{round(s % 60):02d}s"
