"""Coarse LES with the learned closure, and the filtered-DNS pipeline that
trains it (no reference counterpart: an addition to hit3d's solver).

Filtered DNS. A DNS state u on the fine grid is sharp-filtered to |k| <=
kcut and truncated onto the LES grid (`filter_state`); the exact subgrid
stress of that filter is tau_ij = G(u_i u_j) - G(ubar_i ubar_j)
(`subgrid_stress`, evaluated spectrally on both grids: the fine-grid
products are transformed and filtered, the coarse products are free of
aliasing below the cut because 2 kcut < N_les - kcut). The closure input is
the filtered velocity-gradient tensor d_j ubar_i in physical space
(`velocity_gradient`, channel 3 j + i).

LES. d ubar/dt = rhs(ubar) + g S(k) P M [-i k_j tau_hat_ij(net(grad ubar))]
(+ A(ubar) ubar with the HoldResolvedEnergy forcing), with rhs the solver's
Navier-Stokes RHS, the closure term projected and dealiased like the
nonlinear term, g the closure gain and S(k) = exp(tilt (|k|/kc - 1/2)) its
spectral shape; everything restricted to the modes |k| <= kcut. Every piece is a
differentiable torch operation backed by the library: the RHS's backward
is the hand-written adjoint (hit3d_rhs_vjp), the linear spectral operators
use their exact adjoints, and the net's backward runs the BF16-MFMA
backward kernels, so a-posteriori training differentiates whole LES
rollouts.

Gradient convention of the spectral tensors: torch's, dL = Re sum conj(g) dx.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .closure import ClosureNet, apply_net
from .diagnostics import energy_spectrum
from .forcing import NoForcing
from .grid import check_spectral, make_grid, stream, to_physical, to_spectral
from .navierstokes import NSParams

__all__ = ["filter_state", "subgrid_stress", "velocity_gradient", "les_rhs", "les_step",
           "les_rollout", "closure_force", "sphere_mask", "LesStepGraph", "ClosureNet",
           "HoldResolvedEnergy", "hold_energy"]


# ---------------------------------------------------------------------------
# Filtered-DNS pipeline
# ---------------------------------------------------------------------------

def filter_state(u_fine, fine, coarse, kcut, ncomp=3):
    """Sharp spherical filter |k| <= kcut of a fine-grid spectral field,
    expressed on the coarse grid (scaled to the coarse grid's rfftn)."""
    out = torch.empty((ncomp, coarse.Nx, coarse.Ny, coarse.Nz // 2 + 1),
                      dtype=coarse.complex_dtype, device=coarse.device)
    _lib.call("hit3d_spectral_truncate", fine.desc, u_fine.contiguous().data_ptr(), coarse.desc,
              out.data_ptr(), ncomp, float(kcut), stream())
    return out


def _products(u_phys, grid):
    uu = torch.empty((6, grid.Nx, grid.Ny, grid.Nz), dtype=grid.dtype, device=grid.device)
    _lib.call("hit3d_velocity_products", grid.desc, u_phys.data_ptr(), uu.data_ptr(), stream())
    return uu


def subgrid_stress(u_fine, fine, coarse, kcut):
    """(ubar, tau_hat): the filtered state on the coarse grid and the
    spectral subgrid stress tau_ij = G(u_i u_j) - G(ubar_i ubar_j) (6
    components, order 00, 01, 02, 11, 12, 22) on the coarse grid."""
    ubar = filter_state(u_fine, fine, coarse, kcut)
    uu_f = to_spectral(_products(to_physical(u_fine, fine), fine), fine)
    g_uu = filter_state(uu_f, fine, coarse, kcut, ncomp=6)
    del uu_f
    vv = to_spectral(_products(to_physical(ubar, coarse), coarse), coarse)
    g_vv = filter_state(vv, coarse, coarse, kcut, ncomp=6)
    return ubar, g_uu - g_vv


def velocity_gradient(ubar, grid):
    """Physical d_j ubar_i, channel 3 j + i, fp32/fp64 [9][Nx][Ny][Nz]."""
    G = _Grad9.apply(ubar, grid)
    return _ToPhysical.apply(G, grid)


# ---------------------------------------------------------------------------
# Differentiable building blocks (library kernels, hand-written adjoints)
# ---------------------------------------------------------------------------

def _deriv(x, grid, op):
    out = torch.empty_like(x)
    _lib.call("hit3d_spectral_derivative", grid.desc, x.data_ptr(), out.data_ptr(), x.shape[0], op,
              stream())
    return out


class _Grad9(torch.autograd.Function):
    """u[3] -> G[3 j + i] = i k_j u_i; adjoint: sum_j conj(i k_j) g = -i k_j g."""

    @staticmethod
    def forward(ctx, u, grid):
        ctx.grid = grid
        u = u.contiguous()
        out = torch.empty((9,) + tuple(u.shape[1:]), dtype=u.dtype, device=u.device)
        for j in range(3):
            _lib.call("hit3d_spectral_derivative", grid.desc, u.data_ptr(), out[3 * j].data_ptr(), 3, j,
                      stream())
        return out

    @staticmethod
    def backward(ctx, g):
        g = g.contiguous()
        out = sum(_deriv(g[3 * j:3 * j + 3].contiguous(), ctx.grid, j) for j in range(3))
        return -out, None


class _ToPhysical(torch.autograd.Function):
    """irfftn with 1/N; adjoint (w / N) rfftn(g)."""

    @staticmethod
    def forward(ctx, x, grid):
        ctx.grid = grid
        return to_physical(x.contiguous(), grid)

    @staticmethod
    def backward(ctx, g):
        grid = ctx.grid
        n = float(grid.Nx * grid.Ny * grid.Nz)
        return to_spectral(g.contiguous(), grid) * (grid.weights / n), None


class _ToSpectral(torch.autograd.Function):
    """rfftn (unnormalised); adjoint N irfftn(g / w)."""

    @staticmethod
    def forward(ctx, x, grid):
        ctx.grid = grid
        return to_spectral(x.contiguous(), grid)

    @staticmethod
    def backward(ctx, g):
        grid = ctx.grid
        n = float(grid.Nx * grid.Ny * grid.Nz)
        return to_physical((g / grid.weights).contiguous(), grid) * n, None


class _StressDiv(torch.autograd.Function):
    """tau_hat[6] -> F_i = -sum_j i k_j tau_ij; adjoint by the library kernel."""

    @staticmethod
    def forward(ctx, t, grid):
        ctx.grid = grid
        out = torch.empty((3,) + tuple(t.shape[1:]), dtype=t.dtype, device=t.device)
        _lib.call("hit3d_stress_divergence", grid.desc, t.contiguous().data_ptr(), out.data_ptr(), 0,
                  stream())
        return out

    @staticmethod
    def backward(ctx, g):
        out = torch.empty((6,) + tuple(g.shape[1:]), dtype=g.dtype, device=g.device)
        _lib.call("hit3d_stress_divergence", ctx.grid.desc, g.contiguous().data_ptr(), out.data_ptr(),
                  1, stream())
        return out, None


def _dealias_project_raw(x, grid):
    out = torch.empty_like(x)
    _lib.call("hit3d_dealias_project", grid.desc, x.contiguous().data_ptr(), out.data_ptr(), 3, 1, 1,
              stream())
    return out


class _DealiasProject(torch.autograd.Function):
    """P M per mode: real and symmetric, hence its own adjoint."""

    @staticmethod
    def forward(ctx, x, grid):
        ctx.grid = grid
        return _dealias_project_raw(x, grid)

    @staticmethod
    def backward(ctx, g):
        return _dealias_project_raw(g, ctx.grid), None


class _NSRhs(torch.autograd.Function):
    """The solver's RHS (hit3d_rhs) with the hand-written adjoint
    (hit3d_rhs_vjp) as its backward."""

    @staticmethod
    def forward(ctx, u, rhs, params, t):
        ctx.save_for_backward(u)
        ctx.rhs, ctx.params, ctx.t = rhs, params, t
        return rhs(u.detach(), params, t, dealiased=True)   # LES states stay inside the 2/3 box

    @staticmethod
    def backward(ctx, g):
        (u,) = ctx.saved_tensors
        from .adjoint import make_ns_rhs_vjp
        return make_ns_rhs_vjp(ctx.rhs.grid)(u, g.contiguous(), ctx.params, ctx.t, dealiased=True), \
            None, None, None


# ---------------------------------------------------------------------------
# LES right-hand side and time stepping
# ---------------------------------------------------------------------------

_PAIRS = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))


def clip_backscatter(tau, g):
    """tau (6 x V) set to 0 where the modelled transfer -tau_ij S_ij is
    negative; S_ij = (d_j u_i + d_i u_j)/2 from g (channel 3 j + i). Torch
    statement of what hit3d_stress_unpack_clip does inside the net."""
    pi = 0.0
    for s, (i, j) in enumerate(_PAIRS):
        sij = 0.5 * (g[3 * j + i] + g[3 * i + j])
        pi = pi - (1.0 if i == j else 2.0) * tau[s] * sij
    return tau * (pi >= 0).to(tau.dtype)


def closure_force(u, grid, net):
    """gain * S(k) * P M [-i k_j tau_hat_ij], tau = net(grad u), with
    backscatter clipping when the net asks for it and S the net's spectral
    shape exp(tilt (|k|/kc - 1/2)) (differentiable)."""
    if grid.dtype != torch.float32:
        raise ValueError("the closure runs on float32 LES grids")
    g = velocity_gradient(u, grid)
    V = grid.Nx * grid.Ny * grid.Nz
    tau = apply_net(net, g.reshape(9, V)).reshape(6, grid.Nx, grid.Ny, grid.Nz)  # clipping inside
    f = _DealiasProject.apply(_StressDiv.apply(_ToSpectral.apply(tau, grid), grid), grid)
    return net.gain * (net.spectral_shape() * f)


def sphere_mask(grid, kcut):
    """0/1 mask of the modes with |k| <= kcut: the LES resolves exactly the
    sharp spherical filter's modes (the solver's 2/3 box is a cube, whose
    corners reach sqrt(3) kcut). The tensor carries `kcut`, which lets the
    fused library kernel apply it."""
    m = (grid.k2 <= float(kcut) ** 2).to(grid.dtype)
    m.kcut = float(kcut)
    return m


def _les_rhs_fused(u, rhs, p, hold, net, t, kcut, defer_hold=False):
    """les_rhs without autograd: the NS RHS, then the closure force, the
    sum and the spherical mask in one library pass (hit3d_les_combine),
    then the energy hold. defer_hold: return (RHS without the hold's term,
    its device coefficients) for hit3d_rk4_combine to add A u."""
    grid = rhs.grid
    du = rhs(u, p, t, dealiased=True)
    out = torch.empty_like(du)
    if net is not None:
        V = grid.Nx * grid.Ny * grid.Nz
        tau = net(velocity_gradient(u, grid).reshape(9, V)).reshape(6, grid.Nx, grid.Ny, grid.Nz)
        tau_hat = to_spectral(tau, grid)
        _lib.call("hit3d_les_combine", grid.desc, du.data_ptr(), tau_hat.data_ptr(),
                  net.spectral_shape().data_ptr(), net.gain.data_ptr(), kcut, out.data_ptr(), stream())
    else:
        _lib.call("hit3d_les_combine", grid.desc, du.data_ptr(), None, None, None, kcut, out.data_ptr(),
                  stream())
    if hold is not None and defer_hold:
        return out, _hold_raw(u, out, grid, hold.energy, hold.tau, apply=False)[1]
    if hold is not None:
        out = _hold_raw(u, out, grid, hold.energy, hold.tau)[0]
    return out


@dataclass(frozen=True)
class HoldResolvedEnergy:
    """LES forcing f = A(u) u that holds the resolved energy: A is set from
    every other term of the LES RHS (closure included) so that
    dE/dt = -(E - energy)/tau, or dE/dt = 0 with tau=None. It is the LES
    counterpart of the DNS's constant-energy forcing (forcing.py:5-20 for
    the form f = A u): the DNS holds its total energy, the LES holds its
    resolved energy at the filtered DNS's mean. Only les_rhs knows it; the
    solver's own RHS runs unforced underneath (hit3d_hold_energy)."""

    energy: float
    tau: float | None = 1.0

    @property
    def label(self):
        return f"HoldResolvedEnergy_E{float(self.energy):g}_tau{self.tau}"


def _hold_raw(u, du, grid, energy, tau, apply=True):
    u = check_spectral(u, grid, 3)
    if du.shape != u.shape or du.dtype != u.dtype or du.device != u.device:
        raise ValueError("du must match u in shape, dtype and device")
    out = torch.empty_like(du) if apply else None
    coef = torch.empty(3, dtype=torch.float64, device=du.device)
    ws = grid.workspace(_lib.OP_REDUCE)
    _lib.call("hit3d_hold_energy", grid.desc, u.contiguous().data_ptr(), du.contiguous().data_ptr(),
              None if out is None else out.data_ptr(), float(energy), 0.0 if tau is None else float(tau),
              coef.data_ptr(), ws.data_ptr(), ws.numel(), stream())
    return out, coef


class _HoldEnergy(torch.autograd.Function):
    """out = du + A(u, du) u (hit3d_hold_energy); the backward differentiates
    A through its two FP64 sums s1 = sum w Re(conj u du)/Ntot^2 and
    s2 = sum w |u|^2/Ntot^2: dA/ds1 = -1/s2, dA/ds2 = (-r/2 - A)/s2 with
    r = 1/tau (0 without relaxation)."""

    @staticmethod
    def forward(ctx, u, du, grid, energy, tau):
        u, du = u.contiguous(), du.contiguous()
        out, coef = _hold_raw(u, du, grid, energy, tau)
        ctx.save_for_backward(u, du, coef)
        ctx.grid = grid
        ctx.r = 1.0 / float(tau) if tau is not None and 0 < float(tau) < math.inf else 0.0
        return out

    @staticmethod
    def backward(ctx, g):
        u, du, coef = ctx.saved_tensors
        grid = ctx.grid
        A, s2 = coef[0], coef[2]
        tiny = float(torch.finfo(grid.dtype).eps)
        live = (s2 > tiny).to(coef.dtype)               # A = 0 (and constant) when 2E <= tiny
        c = torch.sum(g.real.double() * u.real + g.imag.double() * u.imag)   # Re sum conj(g) u
        safe = torch.where(s2 > 0, s2, torch.ones_like(s2))
        k1 = (live * c * (-1.0 / safe)).to(grid.dtype)
        k2 = (live * c * ((-0.5 * ctx.r - A) / safe)).to(grid.dtype)
        wn = grid.weights / float(grid.Nx * grid.Ny * grid.Nz) ** 2
        gdu = g + (k1 * wn) * u
        gu = A.to(grid.dtype) * g + (k1 * wn) * du + (2.0 * k2 * wn) * u
        return gu, gdu, None, None, None


def hold_energy(u, du, grid, energy, tau=1.0):
    """du + A u with A from hit3d_hold_energy (differentiable in u and du)."""
    return _HoldEnergy.apply(u, du, grid, energy, tau)


def les_rhs(u, rhs, params, net, t=0.0, mask=None):
    """NS RHS + closure force, restricted to `mask`; with a HoldResolvedEnergy
    forcing the NS part runs unforced and the energy-holding term is added
    last, from the masked total."""
    hold = params.forcing if isinstance(params.forcing, HoldResolvedEnergy) else None
    p = NSParams(nu=params.nu, forcing=NoForcing()) if hold is not None else params
    if not torch.is_grad_enabled() and rhs.grid.dtype == torch.float32 and (
            mask is None or getattr(mask, "kcut", None) is not None):
        return _les_rhs_fused(u, rhs, p, hold, net, t, math.inf if mask is None else mask.kcut)
    du = _NSRhs.apply(u, rhs, p, t)
    if net is not None:
        du = du + closure_force(u, rhs.grid, net)
    if mask is not None:
        du = du * mask
    if hold is not None:
        du = hold_energy(u, du, rhs.grid, hold.energy, hold.tau)
    return du


def les_step(u, rhs, params, net, dt, t=0.0, mask=None, out=None):
    """Classic RK4 (integrators.py:32-38 structure) on the LES RHS. Without
    gradients each stage's sums are one library pass (hit3d_rk4_combine), and
    `out` may name the array that receives the new state (u itself
    included); with gradients they are tensor ops and `out` must be None."""
    if not torch.is_grad_enabled():
        return _les_step_fused(u, rhs, params, net, dt, t, mask, out)
    if out is not None:
        raise ValueError("out= is only taken by steps without gradients")
    k1 = les_rhs(u, rhs, params, net, t, mask)
    k2 = les_rhs(torch.add(u, k1, alpha=0.5 * dt), rhs, params, net, t + 0.5 * dt, mask)
    acc = torch.add(u, k1, alpha=dt / 6.0).add(k2, alpha=dt / 3.0)
    k3 = les_rhs(torch.add(u, k2, alpha=0.5 * dt), rhs, params, net, t + 0.5 * dt, mask)
    acc = acc.add(k3, alpha=dt / 3.0)
    k4 = les_rhs(torch.add(u, k3, alpha=dt), rhs, params, net, t + dt, mask)
    return acc.add(k4, alpha=dt / 6.0)


def _les_step_fused(u, rhs, params, net, dt, t, mask, out):
    grid = rhs.grid
    u = check_spectral(u, grid, 3).contiguous()
    if out is not None and (out.shape != u.shape or out.dtype != u.dtype or out.device != u.device
                            or not out.is_contiguous()):
        raise ValueError("out must be a contiguous array shaped and typed like u")
    us, acc = torch.empty_like(u), torch.empty_like(u)
    new = torch.empty_like(u) if out is None else out
    hold = params.forcing if isinstance(params.forcing, HoldResolvedEnergy) else None
    fused = grid.dtype == torch.float32 and (mask is None or getattr(mask, "kcut", None) is not None)

    def rhs_at(x, tt):
        """(derivative, hold coefficients or None) at the stage input x"""
        if hold is not None and fused:
            p = NSParams(nu=params.nu, forcing=NoForcing())
            return _les_rhs_fused(x, rhs, p, hold, net, tt, math.inf if mask is None else mask.kcut,
                                  defer_hold=True)
        return les_rhs(x, rhs, params, net, tt, mask), None

    def combine(stage, x, tt, dst):
        k, coef = rhs_at(x, tt)
        _lib.call("hit3d_rk4_combine", grid.desc, stage, float(dt), u.data_ptr(), k.data_ptr(),
                  None if coef is None else coef.data_ptr(), None if coef is None else x.data_ptr(),
                  acc.data_ptr(), dst.data_ptr(), stream())

    combine(1, u, t, us)
    combine(2, us, t + 0.5 * dt, us)
    combine(3, us, t + 0.5 * dt, us)
    combine(4, us, t + dt, new)   # reads acc, k and us only
    return new


class LesStepGraph:
    """One LES step without gradients, captured as a HIP graph and replayed
    (the 64^3 step is a hundred small launches; replaying the graph removes
    their host cost). The captured kernels issue no memsets and no host
    synchronisation; workspaces are allocated by the warm-up steps."""

    def __init__(self, u, rhs, params, net, dt, mask=None):
        check_les_state(u, rhs.grid)
        self.u = u.clone()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.no_grad(), torch.cuda.stream(side):
            w = self.u.clone()
            for _ in range(2):
                w = les_step(w, rhs, params, net, dt, 0.0, mask)
        torch.cuda.current_stream().wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        with torch.no_grad(), torch.cuda.graph(self.graph):
            les_step(self.u, rhs, params, net, dt, 0.0, mask, out=self.u)   # new state over the old

    def step(self):
        self.graph.replay()
        return self.u


def check_les_state(u, grid):
    """LES states must lie inside the 2/3 box (the RHS runs pruned)."""
    from .navierstokes import is_dealiased
    if not is_dealiased(check_spectral(u, grid, 3), grid):
        raise ValueError("LES state has content outside the 2/3 box; filter it first")


def les_rollout(u, rhs, params, net, dt, nsteps, t0=0.0, observe=None, every=1, mask=None,
                graph=True):
    """nsteps LES steps without gradients (replayed from a HIP graph unless
    graph=False); observe(n, u) every `every` steps."""
    check_les_state(u, rhs.grid)
    if graph and nsteps > 0:
        g = LesStepGraph(u, rhs, params, net, dt, mask)
        for n in range(int(nsteps)):
            u = g.step()
            if observe is not None and (n + 1) % every == 0:
                observe(n + 1, u)
        return u.clone()
    with torch.no_grad():
        for n in range(int(nsteps)):
            u = les_step(u, rhs, params, net, dt, t0 + n * dt, mask)
            if observe is not None and (n + 1) % every == 0:
                observe(n + 1, u)
    return u
