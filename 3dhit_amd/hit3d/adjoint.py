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

The physical scalars are differentiable too, as NSParams' leaves are in the
reference (navierstokes.py:1-8, tests/test_grad.py:43-45, 59-74): every VJP
here can accumulate [dJ/dnu, dJ/dA (LinearForcing), dJ/deps (BandForcing)]
into a device float64 vector (`param_grad`), reduced on the GPU in FP64 from
the stage cotangents; `differentiable_rollout` returns them to torch
autograd when nu, A or eps are tensors that require grad.
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from .forcing import BandForcing, LinearForcing
from .grid import DEALIASED_MARK, check_spectral, stream
from .navierstokes import known_dealiased
from .integrators import rollout as _rollout

__all__ = ["make_ns_rhs_vjp", "rk4_step_vjp", "rollout_vjp", "differentiable_rollout",
           "param_grad_buffer"]


def make_ns_rhs_vjp(grid):
    """v -> (d rhs / d u_hat)^* v at u_hat, for `rhs = make_ns_rhs(grid)`."""

    def vjp(u_hat, v, params, t, param_grad=None, dealiased=None):
        """`param_grad` (device float64, 3 entries): the derivatives of
        Re<v, rhs(u_hat)> in nu, A, eps are added to it. The pruned inverse
        transforms are taken when u_hat carries the dealiased mark or
        `dealiased` says so (no device read)."""
        u = check_spectral(u_hat, grid, 3)
        vv = check_spectral(v, grid, 3, "v")
        out = torch.empty_like(u)
        ws = grid.workspace(_lib.OP_RHS_VJP)
        f = params.forcing.desc()
        if dealiased is None:
            dealiased = getattr(u, DEALIASED_MARK, None) == u._version
        _lib.call("hit3d_rhs_vjp", grid.desc, u.data_ptr(), vv.data_ptr(), out.data_ptr(),
                  float(params.nu), ctypes.byref(f), float(t), 1 if dealiased else 0,
                  _pg_ptr(param_grad, u), ws.data_ptr(), ws.numel(), stream())
        return out

    return vjp


def param_grad_buffer(device):
    """Zeroed device accumulator [dJ/dnu, dJ/dA, dJ/deps] (float64)."""
    return torch.zeros(3, dtype=torch.float64, device=device)


def _pg_ptr(pg, like):
    if pg is None:
        return None
    if pg.dtype != torch.float64 or pg.numel() != 3 or pg.device != like.device or not pg.is_contiguous():
        raise ValueError("param_grad must be a contiguous float64 device tensor of 3 entries")
    return pg.data_ptr()


def _step_vjp_into(grid, u, ub, params, dt, t, out, param_grad=None, stages=None, dealiased=None):
    """Without `stages` the step's stage inputs are recomputed. The pruned
    transforms (stage recomputation and the inverse transforms of the
    stage adjoints) are taken when u carries the dealiased mark (rollout's
    checkpoints do) or `dealiased` says so: no device read."""
    ws = grid.workspace(_lib.OP_RK4_VJP)
    f = params.forcing.desc()
    if dealiased is None:
        dealiased = getattr(u, DEALIASED_MARK, None) == u._version
    inside = 1 if dealiased else 0
    _lib.call("hit3d_rk4_step_vjp", grid.desc, u.data_ptr(), ub.data_ptr(), out.data_ptr(),
              float(params.nu), float(dt), float(t), ctypes.byref(f), inside,
              None if stages is None else stages.data_ptr(), _pg_ptr(param_grad, u), ws.data_ptr(),
              ws.numel(), stream())


def rk4_step_vjp(rhs, u_hat, ubar, params, dt, t, param_grad=None):
    """Cotangent of u_hat given the cotangent `ubar` of rk4_step's result
    (parameter derivatives added to `param_grad` when given)."""
    g = rhs.grid
    u = check_spectral(u_hat, g, 3)
    ub = check_spectral(ubar, g, 3, "ubar")
    out = torch.empty_like(u)
    _step_vjp_into(g, u, ub, params, dt, t, out, param_grad)
    return out


# share of the free device memory a segment's record may take
_RECORD_SHARE = 0.5


def _record_segment(rhs, u0, params, dt, nsteps, t0):
    """([nsteps][4][3][Nx][Ny][Nz/2+1], dealiased): per step its start state
    and the inputs of stages 2..4 (hit3d_rk4_steps_record), and whether they
    lie inside the 2/3 box; (None, None) when the record does not fit in
    device memory."""
    g = rhs.grid
    per_step = 4 * u0.numel() * u0.element_size()
    free, _ = torch.cuda.mem_get_info(u0.device)
    if nsteps * per_step > _RECORD_SHARE * free:
        return None, None
    rec = torch.empty((nsteps, 4) + tuple(u0.shape), dtype=u0.dtype, device=u0.device)
    inside = known_dealiased(u0, g)
    u = u0.clone()
    pruned = 1 if inside else 0
    ws = g.workspace(_lib.OP_RK4, 3, pruned)
    f = params.forcing.desc()
    _lib.call("hit3d_rk4_steps_record", g.desc, u.data_ptr(), float(params.nu), float(dt), int(nsteps),
              float(t0), ctypes.byref(f), pruned, rec.data_ptr(), ws.data_ptr(), ws.numel(), stream())
    return rec, inside


def rollout_vjp(rhs, u_hat, ubar, params, dt, nsteps, t0=0.0, checkpoint=20, param_grad=None,
                checkpoints=None):
    """Cotangent of rollout's input given the cotangent of its output, with
    checkpointed recomputation: every segment of `checkpoint` steps is run
    again once, keeping each step's start state and stage inputs (when that
    record fits in half the free device memory; otherwise the states only,
    and each step's adjoint recomputes its stage inputs). Memory:
    nsteps/checkpoint states plus one segment's record. `checkpoints` maps
    segment starts to states the forward pass already kept (as
    differentiable_rollout does); missing ones are recomputed. The
    parameter derivatives of every step are added to `param_grad` when
    given."""
    g = rhs.grid
    if checkpoint < 1:
        raise ValueError("checkpoint must be >= 1")
    u0 = check_spectral(u_hat, g, 3)
    ub = check_spectral(ubar, g, 3, "ubar").clone()
    starts = list(range(0, int(nsteps), int(checkpoint)))
    ckpt = dict(checkpoints or {})
    ckpt[0] = u0
    for a, b in zip(starts[:-1], starts[1:]):
        if b not in ckpt:
            ckpt[b] = _rollout(rhs, ckpt[a], params, dt, b - a, t0 + a * dt)
    for s in reversed(starts):
        e = min(s + checkpoint, int(nsteps))
        rec, inside = _record_segment(rhs, ckpt[s], params, dt, e - s, t0 + s * dt)
        if rec is not None:
            for n in range(e - 1, s - 1, -1):
                _step_vjp_into(g, rec[n - s, 0], ub, params, dt, t0 + n * dt, ub, param_grad,
                               stages=rec[n - s, 1], dealiased=inside)
            del rec
            continue
        states = [ckpt[s]]
        for n in range(s, e - 1):
            states.append(_rollout(rhs, states[-1], params, dt, 1, t0 + n * dt))
        for n in range(e - 1, s - 1, -1):
            _step_vjp_into(g, states[n - s], ub, params, dt, t0 + n * dt, ub, param_grad)
    return ub


def _leaf(x):
    return x if isinstance(x, torch.Tensor) else None


class _Rollout(torch.autograd.Function):
    """rollout with the hand-written adjoint as its backward (the role of
    jax.custom_vjp around the reference's scan). nu and the forcing
    coefficient enter as tensors when they are to be differentiated."""

    @staticmethod
    def forward(ctx, u_hat, nu, coef, rhs, params, dt, nsteps, t0, checkpoint):
        # the forward runs segment by segment and keeps each segment's start
        # state: the backward's checkpoints, not computed twice
        u = u_hat.detach()
        keep = {}
        for s in range(0, nsteps, checkpoint):
            if s > 0:
                keep[s] = u
            u = _rollout(rhs, u, params, dt, min(checkpoint, nsteps - s), t0 + s * dt)
        ctx.save_for_backward(u_hat)
        ctx.cfg = (rhs, params, dt, nsteps, t0, checkpoint)
        ctx.keep = keep
        return u if nsteps > 0 else u.clone()

    @staticmethod
    def backward(ctx, grad_out):
        (u0,) = ctx.saved_tensors
        rhs, params, dt, nsteps, t0, checkpoint = ctx.cfg
        need_nu, need_coef = ctx.needs_input_grad[1], ctx.needs_input_grad[2]
        pg = param_grad_buffer(u0.device) if (need_nu or need_coef) else None
        g = rollout_vjp(rhs, u0, grad_out.contiguous(), params, dt, nsteps, t0, checkpoint, pg,
                        checkpoints=ctx.keep)
        ctx.keep = None
        g_nu = g_coef = None
        if pg is not None:
            host = pg.cpu()
            if need_nu:
                g_nu = host[0].to(params.nu.dtype).reshape(params.nu.shape)
            if need_coef:
                slot, leaf = (1, params.forcing.A) if isinstance(params.forcing, LinearForcing) else \
                             (2, params.forcing.eps)
                g_coef = host[slot].to(leaf.dtype).reshape(leaf.shape)
        return g, g_nu, g_coef, None, None, None, None, None, None


def differentiable_rollout(rhs, u_hat, params, dt, nsteps, t0=0.0, checkpoint=20):
    """rollout(...) whose result back-propagates through the hand-written
    adjoint (use with torch.autograd) to u_hat and, when they are tensors
    that require grad, to params.nu and to the forcing's coefficient
    (LinearForcing.A or BandForcing.eps) — the leaves the reference's
    jax.grad reaches (tests/test_grad.py:43-45, 59-74). Scalar parameters are
    host (CPU) tensors; their gradients come back as such."""
    f = params.forcing
    coef = _leaf(getattr(f, "A", None)) if isinstance(f, LinearForcing) else \
        _leaf(getattr(f, "eps", None)) if isinstance(f, BandForcing) else None
    return _Rollout.apply(u_hat, _leaf(params.nu), coef, rhs, params, dt, int(nsteps), float(t0),
                          int(checkpoint))
