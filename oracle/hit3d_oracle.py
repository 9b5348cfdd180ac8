"""CPU oracle: a float64 NumPy/SciPy restatement of hit3d's pseudo-spectral
Navier–Stokes step.

THIS MODULE IS TEST INFRASTRUCTURE. It is imported only by ``tests/``, by
``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of ``bench.py``,
and there only as the checker (or the timed CPU baseline). The product path
(``3dhit_amd/hit3d``) never imports it and has no CPU fallback.

Every function cites the reference file:line it restates. The reference
computes with jax/jaxlib 0.10.2 (``uv.lock:263-264``); jax is not importable
in this environment, so the restatement is pinned by closed forms and by the
reference's own property tests (ported in ``tests/test_oracle_*.py``):

=====================  ==========================================================
function               pinned by
=====================  ==========================================================
make_grid              closed forms (integer n, ``|n|<=N//3``, weights, shell
                       rule s(s-1) < k2 <= s(s+1)); ``tests/test_grid.py:20-46``
to_physical/spectral   brute-force DFT (N<=8); round trip; Parseval
ddx..laplacian         exact trig derivatives (``tests/test_grid.py:53-81``)
dealias/project        mask closed form; div-free + idempotent (``:84-93``)
rhs                    Beltrami rhs = -nu*u (``tests/test_navierstokes.py:10-19``);
                       brute-force convolution of u x w at N=8
rk4_step/rollout       Beltrami exp(-nu t) decay; RK4 on a scalar linear ODE
                       (amplification polynomial); nu=0 energy conservation
forcings               exact injection (``tests/test_budget.py:39-60``); measured
                       <u.f>; constant-energy invariant dE/dt = 0
energy..spectrum       Parseval, sum(E(k)) = E, Beltrami Omega=E, lambda=sqrt(5)
random_field           target spectrum on shells 2..N//3 (``tests/test_diagnostics.py:37-43``)
philox_uniform         numpy.random.Philox (Philox4x64-10, a library routine)
=====================  ==========================================================

Spec choices that the reference leaves open (identical in the HIP build):

* D2 — constant-energy Lundgren forcing (``LundgrenForcing``): f = A(u)·u with
  A(u) = (eps(u) - (E(u) - E*)/tau) / (2 E(u)), eps = 2 nu Omega, evaluated on
  the state handed to ``term`` (every RK4 stage, like ``BandForcing``), so
  dE/dt = -(E - E*)/tau for the continuous system (0 for tau=None).
* D4 — initial-condition noise: Philox4x64-10 (Salmon et al., SC'11) with key
  (seed, 0); value i of the flat (3, Nx, Ny, Nz) C-order array is word i % 4
  of block counter (i // 4 + 1, 0, 0, 0) — i.e. exactly
  ``numpy.random.Philox(key=[seed, 0], counter=0).random_raw()``; mapped to
  U[-0.5, 0.5) as (x >> 11)·2^-53 - 0.5 (float64) or (x >> 40)·2^-24 - 0.5
  (float32).
"""

from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np
import scipy.fft as sfft

_WORKERS = max(1, len(os.sched_getaffinity(0)))

# Spectrum of the initial fields of the forced/decaying benchmark configs:
# C k^4 exp(-2 (k/4)^2), normalised to E = 0.5 over shells 1..floor(sqrt(3) N//3 + 1/2)
# (for every N >= 64 the omitted tail is below double round-off).
CONFIG_SPECTRUM_C = 0.004155648754181591


# ---------------------------------------------------------------------------
# Grid tables — grid.py:89-132
# ---------------------------------------------------------------------------


@dataclass(frozen=True)
class Grid:
    """Restates ``Grid`` (grid.py:41-86): broadcast-shaped kx (Nx,1,1),
    ky (1,Ny,1), kz (1,1,Nz//2+1); k2, inv_k2, dealias on the spectral grid;
    Hermitian weights (1,1,Nz//2+1)."""

    Nx: int
    Ny: int
    Nz: int
    Lx: float
    Ly: float
    Lz: float
    kx: np.ndarray
    ky: np.ndarray
    kz: np.ndarray
    k2: np.ndarray
    inv_k2: np.ndarray
    dealias: np.ndarray
    weights: np.ndarray

    @property
    def dtype(self):
        return self.k2.dtype


def make_grid(Nx, Ny=None, Nz=None, *, Lx=2 * np.pi, Ly=2 * np.pi,
              Lz=2 * np.pi, dtype=np.float64):
    """grid.py:89-132. Integer frequencies nx=fftfreq(N,1/N) (:101-103),
    k = (2pi/L n).astype(dtype) (:105-107), k2 in dtype (:109), inv_k2 with
    k=0 -> 0 (:110-111), per-axis box mask |n_i| <= N_i//3 and False at k=0
    (:115-120), weights 1 at kz=0 and the even-N Nyquist plane, else 2
    (:122-125)."""
    Ny = Nx if Ny is None else Ny
    Nz = Nx if Nz is None else Nz
    dtype = np.dtype(dtype)
    nx = np.fft.fftfreq(Nx, 1 / Nx)
    ny = np.fft.fftfreq(Ny, 1 / Ny)
    nz = np.arange(Nz // 2 + 1, dtype=float)
    kx = (2 * np.pi / Lx * nx).astype(dtype).reshape(-1, 1, 1)
    ky = (2 * np.pi / Ly * ny).astype(dtype).reshape(1, -1, 1)
    kz = (2 * np.pi / Lz * nz).astype(dtype).reshape(1, 1, -1)
    k2 = kx**2 + ky**2 + kz**2
    with np.errstate(divide="ignore"):
        inv_k2 = np.where(k2 == 0, dtype.type(0), 1 / k2).astype(dtype)
    mask = (
        (np.abs(nx).reshape(-1, 1, 1) <= Nx // 3)
        & (np.abs(ny).reshape(1, -1, 1) <= Ny // 3)
        & (np.abs(nz).reshape(1, 1, -1) <= Nz // 3)
    )
    mask[0, 0, 0] = False
    w = np.full(Nz // 2 + 1, 2, dtype=dtype)
    w[0] = 1
    if Nz % 2 == 0:
        w[-1] = 1
    return Grid(Nx=Nx, Ny=Ny, Nz=Nz, Lx=float(Lx), Ly=float(Ly), Lz=float(Lz),
                kx=kx, ky=ky, kz=kz, k2=k2, inv_k2=inv_k2, dealias=mask,
                weights=w.reshape(1, 1, -1))


def spectral_shape(grid):
    """grid.py:135-137."""
    return (3, grid.Nx, grid.Ny, grid.Nz // 2 + 1)


def complex_dtype(grid):
    """fields.py:21-22: complex64 for float32 grids, complex128 for float64."""
    return np.result_type(grid.dtype, np.complex64)


# ---------------------------------------------------------------------------
# Spectral operators — grid.py:146-188
# ---------------------------------------------------------------------------


def ddx(u_hat, grid):
    """grid.py:146-148."""
    return 1j * grid.kx * u_hat


def ddy(u_hat, grid):
    """grid.py:151-153."""
    return 1j * grid.ky * u_hat


def ddz(u_hat, grid):
    """grid.py:156-158."""
    return 1j * grid.kz * u_hat


def laplacian(u_hat, grid):
    """grid.py:161-163."""
    return -grid.k2 * u_hat


def dealias(u_hat, grid):
    """grid.py:166-168."""
    return u_hat * grid.dealias


def project(u_hat, grid):
    """grid.py:171-178: u - k (k.u) inv_k2."""
    div = (grid.kx * u_hat[0] + grid.ky * u_hat[1] + grid.kz * u_hat[2]) * grid.inv_k2
    return u_hat - np.stack((grid.kx * div, grid.ky * div, grid.kz * div))


def to_physical(u_hat, grid):
    """grid.py:181-183: irfftn over axes (1,2,3), 1/N^3 normalisation."""
    out = sfft.irfftn(u_hat, s=(grid.Nx, grid.Ny, grid.Nz), axes=(1, 2, 3),
                      workers=_WORKERS)
    return out.astype(grid.dtype, copy=False)


def to_spectral(u, grid):
    """grid.py:186-188: unnormalised rfftn over axes (1,2,3)."""
    out = sfft.rfftn(u, axes=(1, 2, 3), workers=_WORKERS)
    return out.astype(complex_dtype(grid), copy=False)


# ---------------------------------------------------------------------------
# Forcings — forcing.py:44-163 (+ the constant-energy Lundgren rule, D2)
# ---------------------------------------------------------------------------


def _ntot2(grid):
    """forcing.py:44-45."""
    return float(grid.Nx * grid.Ny * grid.Nz) ** 2


def _weighted_2E(u_hat, grid, mask=None):
    """forcing.py:48-52: 2E restricted to `mask`, Hermitian-weighted."""
    w = grid.weights if mask is None else grid.weights * mask
    return np.sum(w * np.sum(np.abs(u_hat) ** 2, axis=0)) / _ntot2(grid)


def _fmt(x):
    return f"{float(x):g}"


@dataclass(frozen=True)
class NoForcing:
    """forcing.py:55-71."""

    def term(self, u_hat, grid, t):
        return np.zeros((), dtype=u_hat.dtype)

    def injection(self, u_hat, grid):
        return grid.dtype.type(0)

    @property
    def label(self):
        return "NoForcing"


@dataclass(frozen=True)
class LinearForcing:
    """forcing.py:74-106: f = A u, injection 2AE."""

    A: float

    def term(self, u_hat, grid, t):
        return self.A * u_hat

    def injection(self, u_hat, grid):
        return self.A * _weighted_2E(u_hat, grid)

    @property
    def label(self):
        return f"LinearForcing_A{_fmt(self.A)}"


@dataclass(frozen=True)
class BandForcing:
    """forcing.py:109-163: f = eps/(2E_band) u on kmin^2 <= k2 <= kmax^2,
    zero when 2E_band <= finfo(dtype).eps."""

    eps: float
    kmin: float
    kmax: float

    def _mask(self, grid):
        return (self.kmin**2 <= grid.k2) & (grid.k2 <= self.kmax**2)

    def coefficient(self, u_hat, grid):
        twoE = _weighted_2E(u_hat, grid, self._mask(grid))
        tiny = np.finfo(grid.dtype).eps
        return self.eps / twoE if twoE > tiny else 0.0

    def term(self, u_hat, grid, t):
        return self.coefficient(u_hat, grid) * self._mask(grid) * u_hat

    def injection(self, u_hat, grid):
        twoE = _weighted_2E(u_hat, grid, self._mask(grid))
        tiny = np.finfo(grid.dtype).eps
        return grid.dtype.type(self.eps if twoE > tiny else 0.0)

    @property
    def label(self):
        return f"BandForcing_eps{_fmt(self.eps)}_k{_fmt(self.kmin)}-{_fmt(self.kmax)}"


@dataclass(frozen=True)
class LundgrenForcing:
    """Constant-energy linear forcing (no reference symbol; spec delta D2):
    f = A(u) u, A(u) = (2 nu Omega(u) - (E(u) - E*)/tau) / (2 E(u)), zero
    when 2E <= finfo(dtype).eps; tau=None drops the relaxation term."""

    nu: float
    energy: float = 0.5
    tau: float | None = None

    def coefficient(self, u_hat, grid):
        a2 = np.sum(np.abs(u_hat) ** 2, axis=0)
        twoE = np.sum(grid.weights * a2) / _ntot2(grid)
        twoOmega = np.sum(grid.weights * grid.k2 * a2) / _ntot2(grid)
        if not twoE > np.finfo(grid.dtype).eps:
            return 0.0
        rate = self.nu * twoOmega
        if self.tau is not None:
            rate -= (0.5 * twoE - self.energy) / self.tau
        return rate / twoE

    def term(self, u_hat, grid, t):
        return self.coefficient(u_hat, grid) * u_hat

    def injection(self, u_hat, grid):
        return self.coefficient(u_hat, grid) * _weighted_2E(u_hat, grid)

    @property
    def label(self):
        tau = "inf" if self.tau is None else _fmt(self.tau)
        return f"LundgrenForcing_E{_fmt(self.energy)}_tau{tau}"


# ---------------------------------------------------------------------------
# Navier–Stokes RHS — navierstokes.py:62-97
# ---------------------------------------------------------------------------


@dataclass(frozen=True)
class NSParams:
    """navierstokes.py:40-59."""

    nu: float
    forcing: object = NoForcing()


def make_ns_rhs(grid):
    """navierstokes.py:62-97: w = ik x u (:73-77); u, w to physical (:78-79);
    nl = u x w (:82-86); rfftn(nl) + forcing.term (:89); dealias then
    project (:92); minus nu k2 u (:95)."""

    def rhs(u_hat, params, t):
        kx, ky, kz = grid.kx, grid.ky, grid.kz
        w_hat = 1j * np.stack((
            ky * u_hat[2] - kz * u_hat[1],
            kz * u_hat[0] - kx * u_hat[2],
            kx * u_hat[1] - ky * u_hat[0],
        ))
        u = to_physical(u_hat, grid)
        w = to_physical(w_hat, grid)
        nl = np.stack((
            u[1] * w[2] - u[2] * w[1],
            u[2] * w[0] - u[0] * w[2],
            u[0] * w[1] - u[1] * w[0],
        ))
        du = to_spectral(nl, grid) + params.forcing.term(u_hat, grid, t)
        du = project(dealias(du, grid), grid)
        return (du - params.nu * grid.k2 * u_hat).astype(u_hat.dtype, copy=False)

    rhs.grid = grid
    return rhs


# ---------------------------------------------------------------------------
# Time integration — integrators.py:32-53
# ---------------------------------------------------------------------------


def rk4_step(rhs, u_hat, params, dt, t):
    """integrators.py:32-38 (classic RK4)."""
    k1 = rhs(u_hat, params, t)
    k2 = rhs(u_hat + 0.5 * dt * k1, params, t + 0.5 * dt)
    k3 = rhs(u_hat + 0.5 * dt * k2, params, t + 0.5 * dt)
    k4 = rhs(u_hat + dt * k3, params, t + dt)
    return u_hat + (dt / 6.0) * (k1 + 2 * k2 + 2 * k3 + k4)


def linear_operator(grid, params):
    """navierstokes.py:100-103: -nu k2."""
    return -params.nu * grid.k2


def if_rk4_step(rhs, u_hat, params, dt, t):
    """Integrating-factor (Lawson) RK4 with the hook navierstokes.py:100-103
    (integrators.py:14-16 leaves the scheme open): with L = linear_operator,
    E_h = exp(L dt/2) and N(u) = rhs(u) - L u, classic RK4 on
    v = exp(-L (s - t)) u gives
      u2 = E_h (u + dt/2 k1), u3 = E_h u + dt/2 k2, u4 = E_h^2 u + dt E_h k3,
      u_new = E_h^2 u + dt/6 (E_h^2 k1 + 2 E_h (k2 + k3) + k4),  k_i = N(u_i)."""
    L = linear_operator(rhs.grid, params)
    eh = np.exp(0.5 * dt * L)
    ef = eh * eh
    N = lambda u, tt: rhs(u, params, tt) - L * u  # noqa: E731
    k1 = N(u_hat, t)
    k2 = N(eh * (u_hat + 0.5 * dt * k1), t + 0.5 * dt)
    k3 = N(eh * u_hat + 0.5 * dt * k2, t + 0.5 * dt)
    k4 = N(ef * u_hat + dt * eh * k3, t + dt)
    return ef * u_hat + (dt / 6.0) * (ef * k1 + 2 * eh * (k2 + k3) + k4)


def rollout(rhs, u_hat, params, dt, nsteps, t0=0.0, observe=None, step=None):
    """integrators.py:41-53 (scan over rk4_step, or over `step`, e.g.
    if_rk4_step). `observe(n, u)` is an oracle-only hook called after every
    step (used for spectra cadences)."""
    step = rk4_step if step is None else step
    for i in range(nsteps):
        u_hat = step(rhs, u_hat, params, dt, t0 + i * dt)
        if observe is not None:
            observe(i + 1, u_hat)
    return u_hat


# ---------------------------------------------------------------------------
# Adjoint (vector-Jacobian products) of rhs, rk4_step and rollout. The
# reference differentiates `rollout` with jax.vjp (tests/test_grad.py:29-74);
# this is the same derivative written out by hand. Convention: for a real
# J(u_hat), the gradient g satisfies dJ = Re sum(conj(g) * du_hat) over every
# array entry (torch's convention; jax.vjp's cotangent is conj(g)). Pinned by
# central finite differences (tests/test_oracle.py).
# ---------------------------------------------------------------------------


def _curl(x, grid):
    """K x = i k x x, the map navierstokes.py:73-77 applies to u_hat. K is
    self-adjoint under <a, b> = Re sum(conj(a) b)."""
    kx, ky, kz = grid.kx, grid.ky, grid.kz
    return 1j * np.stack((ky * x[2] - kz * x[1], kz * x[0] - kx * x[2], kx * x[1] - ky * x[0]))


def _cross(a, b):
    return np.stack((a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]))


def _inner(a, b):
    return float(np.sum(np.real(np.conj(a) * b)))


def _forcing_vjp(forcing, u_hat, vp, grid):
    """Cotangent of u_hat through the forcing term, given vp = P M v."""
    if isinstance(forcing, NoForcing):
        return 0.0
    if isinstance(forcing, LinearForcing):
        return forcing.A * vp
    w, nt2 = grid.weights, _ntot2(grid)
    tiny = np.finfo(grid.dtype).eps
    a2 = np.sum(np.abs(u_hat) ** 2, axis=0)
    if isinstance(forcing, BandForcing):
        band = forcing._mask(grid)
        sB = np.sum(w * band * a2) / nt2
        if not sB > tiny:
            return 0.0
        c = forcing.eps / sB
        grad_c = -(c / sB) * 2 * w * band * u_hat / nt2
        return c * band * vp + _inner(band * u_hat, vp) * grad_c
    if isinstance(forcing, LundgrenForcing):
        sE = np.sum(w * a2) / nt2
        sO = np.sum(w * grid.k2 * a2) / nt2
        if not sE > tiny:
            return 0.0
        A = forcing.coefficient(u_hat, grid)
        dA_dE = -A / sE
        if forcing.tau is not None:
            dA_dE -= 0.5 / forcing.tau / sE
        grad_A = (forcing.nu / sE) * 2 * w * grid.k2 * u_hat / nt2 + dA_dE * 2 * w * u_hat / nt2
        return A * vp + _inner(vp, u_hat) * grad_A
    raise TypeError(f"no adjoint for {forcing!r}")


def make_ns_rhs_vjp(grid):
    """VJP of rhs (navierstokes.py:69-95) at the state u_hat: v -> (d rhs)^* v.
    With g = N^3 irfftn(P M v / w) (the adjoint of rfftn), the nonlinear part
    is (w/N^3) rfftn(w_p x g) + K (w/N^3) rfftn(g x u_p)."""
    nt = grid.Nx * grid.Ny * grid.Nz

    def vjp(u_hat, v, params, t):
        w = grid.weights
        vp = project(dealias(v, grid), grid)
        g = nt * to_physical(vp / w, grid)
        u = to_physical(u_hat, grid)
        om = to_physical(_curl(u_hat, grid), grid)
        a_hat = (w / nt) * to_spectral(_cross(om, g), grid)
        b_hat = (w / nt) * to_spectral(_cross(g, u), grid)
        out = a_hat + _curl(b_hat, grid) - params.nu * grid.k2 * v
        return out + _forcing_vjp(params.forcing, u_hat, vp, grid)

    return vjp


def rhs_param_vjp(u_hat, v, params, grid):
    """Derivatives of Re<v, rhs(u_hat)> with respect to the physical scalars
    NSParams carries (navierstokes.py:40-59, forcing.py:74-163; the leaves
    jax.grad reaches in tests/test_grad.py:43-45, 59-74): [d/dnu, d/dA of
    LinearForcing (else 0), d/deps of BandForcing (else 0)]. rhs depends on
    nu through -nu k2 u, on A through P M (A u), on eps through
    P M (eps / 2E_band(u) band u) (0 below the finfo(dtype).eps guard)."""
    out = np.zeros(3)
    out[0] = _inner(v, -grid.k2 * u_hat)
    f = params.forcing
    if isinstance(f, LinearForcing):
        out[1] = _inner(v, project(dealias(u_hat, grid), grid))
    elif isinstance(f, BandForcing):
        band = f._mask(grid)
        sB = _weighted_2E(u_hat, grid, band)
        if sB > np.finfo(grid.dtype).eps:
            out[2] = _inner(v, project(dealias(band * u_hat, grid), grid)) / sB
    return out


def rk4_step_vjp(rhs, rhs_vjp, u_hat, ubar, params, dt, t, pgrad=None):
    """VJP of rk4_step (integrators.py:32-38): recomputes the stage inputs,
    then runs the four stage adjoints backwards. `pgrad` (array of 3)
    accumulates the parameter derivatives, rhs_param_vjp of every stage
    input with its stage cotangent."""
    k1 = rhs(u_hat, params, t)
    y2 = u_hat + 0.5 * dt * k1
    k2 = rhs(y2, params, t + 0.5 * dt)
    y3 = u_hat + 0.5 * dt * k2
    k3 = rhs(y3, params, t + 0.5 * dt)
    y4 = u_hat + dt * k3
    out = ubar.copy()
    kb1 = kb4 = (dt / 6.0) * ubar
    kb2 = kb3 = (dt / 3.0) * ubar
    grid = rhs.grid

    def stage(y, kb, tt):
        if pgrad is not None:
            pgrad[:] = pgrad + rhs_param_vjp(y, kb, params, grid)
        return rhs_vjp(y, kb, params, tt)
    yb = stage(y4, kb4, t + dt)
    out, kb3 = out + yb, kb3 + dt * yb
    yb = stage(y3, kb3, t + 0.5 * dt)
    out, kb2 = out + yb, kb2 + 0.5 * dt * yb
    yb = stage(y2, kb2, t + 0.5 * dt)
    out, kb1 = out + yb, kb1 + 0.5 * dt * yb
    return out + stage(u_hat, kb1, t)


def rollout_vjp(rhs, rhs_vjp, u_hat, ubar, params, dt, nsteps, t0=0.0, checkpoint=20, pgrad=None):
    """VJP of rollout (integrators.py:41-53) with checkpointed recomputation:
    the forward keeps every `checkpoint`-th state; the backward recomputes
    each segment's states from its checkpoint and runs the step adjoints in
    reverse (parameter derivatives accumulated into `pgrad`, if given)."""
    ckpt = {0: u_hat}
    u = u_hat
    for n in range(nsteps):
        u = rk4_step(rhs, u, params, dt, t0 + n * dt)
        if (n + 1) % checkpoint == 0 and n + 1 < nsteps:
            ckpt[n + 1] = u
    for start in sorted(ckpt, reverse=True):
        stop = min(start + checkpoint, nsteps)
        states = [ckpt[start]]
        for n in range(start, stop - 1):
            states.append(rk4_step(rhs, states[-1], params, dt, t0 + n * dt))
        for n in range(stop - 1, start - 1, -1):
            ubar = rk4_step_vjp(rhs, rhs_vjp, states[n - start], ubar, params, dt, t0 + n * dt, pgrad)
    return ubar


# ---------------------------------------------------------------------------
# Diagnostics — diagnostics.py:45-114, 162-172
# ---------------------------------------------------------------------------


def _weighted_sum(a2, grid, extra=None):
    """diagnostics.py:49-53."""
    w = grid.weights if extra is None else grid.weights * extra
    return np.sum(w * a2) / _ntot2(grid)


def energy(u_hat, grid):
    """diagnostics.py:56-58."""
    return _weighted_sum(np.sum(np.abs(u_hat) ** 2, axis=0), grid) / 2


def enstrophy(u_hat, grid):
    """diagnostics.py:61-65."""
    return _weighted_sum(np.sum(np.abs(u_hat) ** 2, axis=0), grid, extra=grid.k2) / 2


def dissipation(u_hat, grid, nu):
    """diagnostics.py:68-70."""
    return 2 * nu * enstrophy(u_hat, grid)


def shells(grid):
    """diagnostics.py:76-82: shell = floor(sqrt(k2)/k0 + 1/2) in the grid
    dtype, int32; smax = ceil(sqrt(3) max(N)/2) + 1."""
    k0 = 2 * np.pi / grid.Lx
    smax = int(np.ceil(np.sqrt(3) * max(grid.Nx, grid.Ny, grid.Nz) / 2)) + 1
    dt = grid.dtype.type
    shell = np.floor(np.sqrt(grid.k2) / dt(k0) + dt(0.5)).astype(np.int32)
    return shell, smax, k0


def _shell_sum(a2, grid):
    """diagnostics.py:85-91: bincount of weights*a2/(2 Ntot^2), bin 0 dropped."""
    shell, smax, _ = shells(grid)
    e = (grid.weights * a2) / (2 * _ntot2(grid))
    return np.bincount(shell.ravel(), weights=e.ravel(), minlength=smax + 1)[1:smax + 1]


def energy_spectrum(u_hat, grid):
    """diagnostics.py:94-101: (ks, E) with E[s-1] the energy in shell s."""
    _, smax, k0 = shells(grid)
    ks = np.arange(1, smax + 1) * k0
    return ks, _shell_sum(np.sum(np.abs(u_hat) ** 2, axis=0), grid)


def component_spectra(u_hat, grid):
    """diagnostics.py:104-114."""
    _, smax, k0 = shells(grid)
    ks = np.arange(1, smax + 1) * k0
    return (ks, *[_shell_sum(np.abs(u_hat[i]) ** 2, grid) for i in range(3)])


def energy_budget(u_hat, grid, nu, forcing):
    """diagnostics.py:162-172."""
    return {"E": energy(u_hat, grid), "eps": dissipation(u_hat, grid, nu),
            "P": forcing.injection(u_hat, grid)}


# ---------------------------------------------------------------------------
# Initial conditions — fields.py:25-87 (noise per spec delta D4)
# ---------------------------------------------------------------------------


def philox_uniform(seed, shape, dtype):
    """U[-0.5, 0.5) from Philox4x64-10, key (seed, 0), block counter
    i//4 + 1 (numpy.random.Philox's own stream). Restates the noise draw of
    fields.py:45-46 with a generator both this oracle and the GPU can run."""
    n = int(np.prod(shape))
    raw = np.random.Philox(key=[int(seed), 0], counter=0).random_raw(n)
    dtype = np.dtype(dtype)
    if dtype == np.float32:
        u = (raw >> np.uint64(40)).astype(np.float32) * np.float32(2.0**-24)
    else:
        u = (raw >> np.uint64(11)).astype(np.float64) * 2.0**-53
    return (u - dtype.type(0.5)).astype(dtype).reshape(shape)


def default_spectrum(k):
    """fields.py:30-32: k^4 exp(-2 k^2)."""
    return k**4 * np.exp(-2 * k**2)


def config_spectrum(k):
    """C k^4 exp(-2 (k/4)^2), total energy 0.5 on every N >= 64 grid."""
    return CONFIG_SPECTRUM_C * k**4 * np.exp(-2 * (k / 4.0) ** 2)


def random_field(seed, grid, spectrum=default_spectrum):
    """fields.py:35-68: uniform noise -> rfft -> dealias -> project -> per-shell
    rescale c = sqrt(target/E_s) (E_s > 0, target > 0), 0 (target == 0),
    1 (empty shell)."""
    u = philox_uniform(seed, (3, grid.Nx, grid.Ny, grid.Nz), grid.dtype)
    u_hat = project(dealias(to_spectral(u, grid), grid), grid)
    k0 = 2 * np.pi / grid.Lx
    smax = int(np.ceil(np.sqrt(3) * max(grid.Nx, grid.Ny, grid.Nz) / 2)) + 1
    shell, _, _ = shells(grid)
    e_mode = grid.weights * np.sum(np.abs(u_hat) ** 2, axis=0) / (2 * _ntot2(grid))
    Es = np.bincount(shell.ravel(), weights=e_mode.ravel(), minlength=smax + 1)[:smax + 1]
    ks = np.arange(smax + 1, dtype=grid.dtype) * k0
    target = spectrum(ks)
    Es_safe = np.where(Es > 0, Es, 1)
    c = np.where(Es > 0, np.where(target > 0, np.sqrt(target / Es_safe), 0.0), 1.0)
    return (u_hat * c.astype(grid.dtype)[shell]).astype(complex_dtype(grid), copy=False)


def taylor_green(grid):
    """fields.py:71-87."""
    t = grid.dtype
    x = (np.arange(grid.Nx, dtype=t) * (2 * np.pi / grid.Nx)).reshape(-1, 1, 1)
    y = (np.arange(grid.Ny, dtype=t) * (2 * np.pi / grid.Ny)).reshape(1, -1, 1)
    z = (np.arange(grid.Nz, dtype=t) * (2 * np.pi / grid.Nz)).reshape(1, 1, -1)
    u = np.stack((
        np.sin(x) * np.cos(y) * np.cos(z),
        -np.cos(x) * np.sin(y) * np.cos(z),
        np.zeros((grid.Nx, grid.Ny, grid.Nz), dtype=t),
    ))
    return dealias(to_spectral(u, grid), grid)


def beltrami(grid):
    """ABC flow A=B=C=1 (reference tests/conftest.py:16-30): u x w == 0 and
    every active mode has |k|^2 = 1."""
    t = grid.dtype
    x = (np.arange(grid.Nx, dtype=t) * (2 * np.pi / grid.Nx)).reshape(-1, 1, 1)
    y = (np.arange(grid.Ny, dtype=t) * (2 * np.pi / grid.Ny)).reshape(1, -1, 1)
    z = (np.arange(grid.Nz, dtype=t) * (2 * np.pi / grid.Nz)).reshape(1, 1, -1)
    shape = (grid.Nx, grid.Ny, grid.Nz)
    u = np.stack((
        np.broadcast_to(np.sin(z) + np.cos(y), shape),
        np.broadcast_to(np.sin(x) + np.cos(z), shape),
        np.broadcast_to(np.sin(y) + np.cos(x), shape),
    ))
    return to_spectral(u, grid)
