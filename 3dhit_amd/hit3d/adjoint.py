"""Hand-written adjoint of rhs, rk4_step and rollout.

The reference gets these derivatives from jax autodiff through `rollout`
(integrators.py:4-11, tests/test_grad.py:29-74), storing every step. Here the
vector-Jacobian product of the RHS is its own HIP kernel chain
(`hit3d_rhs_vjp`), one RK4 step's adjoint recomputes the stage inputs and
runs the four stage adjoints in reverse (`hit3d_rk4_step_vjp`), and
`rollout_vjp` keeps only every `checkpoint`-th state, recomputing each
segment during the backward sweep.

Gradient convention (torch's): for a real J, the gradient g of a complex
tensor satisfies dJ = Re sum(conj(g) * du) over all entries; jax.vjp's
cotangent for the same J is conj(g).
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from .grid import check_spectral, stream
from .integrators import rollout as _rollout

__all__ = ["make_ns_rhs_vjp", "rk4_step_vjp", "rollout_vjp", "differentiable_rollout"]


def make_ns_rhs_vjp(grid):
    """v -> (d rhs / d u_hat)^* v at u_hat, for `rhs = make_ns_rhs(grid)`."""

    def vjp(u_hat, v, params, t):
        u = check_spectral(u_hat, grid, 3)
        vv = check_spectral(v, grid, 3, "v")
        out = torch.empty_like(u)
        ws = grid.workspace(_lib.OP_RHS_VJP)
        f = params.forcing.desc()
        _lib.call("hit3d_rhs_vjp", grid.desc, u.data_ptr(), vv.data_ptr(), out.data_ptr(),
                  float(params.nu), ctypes.byref(f), float(t), ws.data_ptr(), ws.numel(),
                  stream())
        return out

    return vjp


def _step_vjp_into(grid, u, ub, params, dt, t, out):
    ws = grid.workspace(_lib.OP_RK4_VJP)
    f = params.forcing.desc()
    _lib.call("hit3d_rk4_step_vjp", grid.desc, u.data_ptr(), ub.data_ptr(), out.data_ptr(),
              float(params.nu), float(dt), float(t), ctypes.byref(f), ws.data_ptr(), ws.numel(),
              stream())


def rk4_step_vjp(rhs, u_hat, ubar, params, dt, t):
    """Cotangent of u_hat given the cotangent `ubar` of rk4_step's result."""
    g = rhs.grid
    u = check_spectral(u_hat, g, 3)
    ub = check_spectral(ubar, g, 3, "ubar")
    out = torch.empty_like(u)
    _step_vjp_into(g, u, ub, params, dt, t, out)
    return out


def rollout_vjp(rhs, u_hat, ubar, params, dt, nsteps, t0=0.0, checkpoint=20):
    """Cotangent of rollout's input given the cotangent of its output, with
    checkpointed recomputation (memory: nsteps/checkpoint + checkpoint states)."""
    g = rhs.grid
    if checkpoint < 1:
        raise ValueError("checkpoint must be >= 1")
    u0 = check_spectral(u_hat, g, 3)
    ub = check_spectral(ubar, g, 3, "ubar").clone()
    starts = list(range(0, int(nsteps), int(checkpoint)))
    ckpt = {0: u0}
    for a, b in zip(starts[:-1], starts[1:]):
        ckpt[b] = _rollout(rhs, ckpt[a], params, dt, b - a, t0 + a * dt)
    for s in reversed(starts):
        e = min(s + checkpoint, int(nsteps))
        states = [ckpt[s]]
        for n in range(s, e - 1):
            states.append(_rollout(rhs, states[-1], params, dt, 1, t0 + n * dt))
        for n in range(e - 1, s - 1, -1):
            _step_vjp_into(g, states[n - s], ub, params, dt, t0 + n * dt, ub)
    return ub


class _Rollout(torch.autograd.Function):
    """rollout with the hand-written adjoint as its backward (the role of
    jax.custom_vjp around the reference's scan)."""

    @staticmethod
    def forward(ctx, u_hat, rhs, params, dt, nsteps, t0, checkpoint):
        ctx.save_for_backward(u_hat)
        ctx.cfg = (rhs, params, dt, nsteps, t0, checkpoint)
        return _rollout(rhs, u_hat.detach(), params, dt, nsteps, t0)

    @staticmethod
    def backward(ctx, grad_out):
        (u0,) = ctx.saved_tensors
        rhs, params, dt, nsteps, t0, checkpoint = ctx.cfg
        g = rollout_vjp(rhs, u0, grad_out.contiguous(), params, dt, nsteps, t0, checkpoint)
        return g, None, None, None, None, None, None


def differentiable_rollout(rhs, u_hat, params, dt, nsteps, t0=0.0, checkpoint=20):
    """rollout(...) whose result back-propagates to u_hat through the
    hand-written adjoint (use with torch.autograd)."""
    return _Rollout.apply(u_hat, rhs, params, dt, int(nsteps), float(t0), int(checkpoint))
