"""Pins the CPU oracle (oracle/hit3d_oracle.py) before it is trusted as the
checker: the reference's own property tests (same tolerances, float64),
closed forms, and brute force on tiny grids. Nothing here needs a GPU."""

import numpy as np
import pytest

import hit3d_oracle as O
from conftest import dealiased_random_spectral, fill3

N = 32
NU = 0.05


# --- grid tables: reference tests/test_grid.py:20-46 + closed forms ----------


def test_shapes(grid32):
    g = grid32
    assert g.kx.shape == (N, 1, 1) and g.ky.shape == (1, N, 1)
    assert g.kz.shape == (1, 1, N // 2 + 1)
    assert g.k2.shape == (N, N, N // 2 + 1)
    assert g.inv_k2[0, 0, 0] == 0
    assert g.weights[0, 0, 0] == 1 and g.weights[0, 0, -1] == 1 and g.weights[0, 0, 1] == 2


def test_dealias_mask_edges(grid32):
    m, cut = grid32.dealias, N // 3
    assert not m[0, 0, 0]
    assert m[cut, 0, 0] and not m[cut + 1, 0, 0]
    assert m[0, N - cut, 0] and not m[0, N - cut - 1, 0]
    assert m[0, 0, cut] and not m[0, 0, cut + 1]
    assert m[1, 0, 0]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(64, 64, 64), (16, 8, 4), (32, 16, 64)])
def test_tables_closed_form(dtype, shape):
    Nx, Ny, Nz = shape
    g = O.make_grid(Nx, Ny, Nz, dtype=dtype)
    nx = np.array([i if i < (Nx + 1) // 2 else i - Nx for i in range(Nx)])
    ny = np.array([i if i < (Ny + 1) // 2 else i - Ny for i in range(Ny)])
    nz = np.arange(Nz // 2 + 1)
    n2 = nx[:, None, None] ** 2 + ny[None, :, None] ** 2 + nz[None, None, :] ** 2
    assert np.array_equal(g.k2, n2.astype(dtype))           # k0 = 1 exactly for L = 2 pi
    inv = np.zeros(n2.shape)
    inv[n2 > 0] = 1.0 / n2[n2 > 0]
    assert np.array_equal(g.inv_k2, inv.astype(dtype))     # correctly rounded 1/k2
    box = ((np.abs(nx)[:, None, None] <= Nx // 3) & (np.abs(ny)[None, :, None] <= Ny // 3)
           & (nz[None, None, :] <= Nz // 3))
    box[0, 0, 0] = False
    assert np.array_equal(g.dealias, box)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_shell_rule_integer_form(dtype):
    """floor(sqrt(k2) + 1/2) equals the integer rule s(s-1) < k2 <= s(s+1)
    for every integer k2 up to 3*512^2 (all grids up to 1024^3)."""
    k2 = np.arange(0, 3 * 512 * 512 + 1, dtype=np.int64)
    s_float = np.floor(np.sqrt(k2.astype(dtype)) + dtype(0.5)).astype(np.int64)
    s = np.floor(np.sqrt(k2.astype(np.float64))).astype(np.int64)
    s = np.where(s * (s + 1) < k2, s + 1, s)     # smallest s with k2 <= s(s+1)
    assert np.array_equal(s_float, s)
    g = O.make_grid(16, dtype=dtype)
    shell, smax, k0 = O.shells(g)
    assert smax == int(np.ceil(np.sqrt(3) * 16 / 2)) + 1 and k0 == 1.0
    n2 = g.k2.astype(np.int64)
    ss = np.floor(np.sqrt(n2.astype(np.float64))).astype(np.int64)
    ss = np.where(ss * (ss + 1) < n2, ss + 1, ss)
    assert np.array_equal(shell, ss)


# --- FFT boundary: brute-force DFT, round trip, Parseval --------------------


def _dft_axis(a, axis, sign):
    n = a.shape[axis]
    j = np.arange(n)
    w = np.exp(sign * 2j * np.pi * np.outer(j, j) / n)
    return np.moveaxis(np.tensordot(w, np.moveaxis(a, axis, 0), axes=(1, 0)), 0, axis)


@pytest.mark.parametrize("shape", [(4, 4, 4), (8, 4, 2), (8, 8, 8)])
def test_to_spectral_brute_force(shape):
    g = O.make_grid(*shape)
    u = np.random.default_rng(0).standard_normal((3,) + shape)
    full = _dft_axis(_dft_axis(_dft_axis(u, 1, -1), 2, -1), 3, -1)
    assert np.allclose(O.to_spectral(u, g), full[..., : shape[2] // 2 + 1], atol=1e-12)


@pytest.mark.parametrize("shape", [(4, 4, 4), (8, 4, 2), (8, 8, 8)])
def test_to_physical_brute_force(shape):
    """c2r semantics: complex inverse on x, y, then the z axis uses only the
    real part of the kz=0 and Nyquist bins (Hermitian extension)."""
    Nx, Ny, Nz = shape
    g = O.make_grid(*shape)
    rng = np.random.default_rng(1)
    u_hat = rng.standard_normal((3, Nx, Ny, Nz // 2 + 1)) + 1j * rng.standard_normal(
        (3, Nx, Ny, Nz // 2 + 1))
    a = _dft_axis(_dft_axis(u_hat, 1, +1), 2, +1)
    ext = np.zeros((3, Nx, Ny, Nz), complex)
    ext[..., : Nz // 2 + 1] = a
    ext[..., 0] = a[..., 0].real
    ext[..., Nz // 2] = a[..., Nz // 2].real
    for k in range(1, Nz // 2):
        ext[..., Nz - k] = np.conj(a[..., k])
    ref = _dft_axis(ext, 3, +1).real / (Nx * Ny * Nz)
    assert np.allclose(O.to_physical(u_hat, g), ref, atol=1e-12)


def test_roundtrip_and_anisotropic(grid32):
    u = np.random.default_rng(0).standard_normal((3, N, N, N))
    assert np.allclose(O.to_physical(O.to_spectral(u, grid32), grid32), u)
    g = O.make_grid(16, 8, 4)
    u = np.random.default_rng(2).standard_normal((3, 16, 8, 4))
    assert np.allclose(O.to_physical(O.to_spectral(u, g), g), u)


def test_derivatives_exact_on_trig(grid32):
    """reference tests/test_grid.py:53-81."""
    g = grid32
    x = (np.arange(N) * (2 * np.pi / N)).reshape(-1, 1, 1)
    y = (np.arange(N) * (2 * np.pi / N)).reshape(1, -1, 1)
    z = (np.arange(N) * (2 * np.pi / N)).reshape(1, 1, -1)
    f = np.sin(x) * np.cos(2 * y) * np.sin(3 * z)
    u_hat = O.to_spectral(fill3(f, g), g)
    P = lambda a: O.to_physical(a, g)  # noqa: E731
    assert np.allclose(P(O.ddx(u_hat, g)), fill3(np.cos(x) * np.cos(2 * y) * np.sin(3 * z), g), atol=1e-10)
    assert np.allclose(P(O.ddy(u_hat, g)), fill3(-2 * np.sin(x) * np.sin(2 * y) * np.sin(3 * z), g), atol=1e-10)
    assert np.allclose(P(O.ddz(u_hat, g)), fill3(3 * np.sin(x) * np.cos(2 * y) * np.cos(3 * z), g), atol=1e-10)
    assert np.allclose(P(O.laplacian(u_hat, g)), fill3(-14 * f, g), atol=1e-9)


def test_projection_divergence_free_and_idempotent(grid32):
    """reference tests/test_grid.py:84-93."""
    g = grid32
    rng = np.random.default_rng(1)
    s = (3, N, N, N // 2 + 1)
    u_hat = O.project(rng.standard_normal(s) + 1j * rng.standard_normal(s), g)
    div = g.kx * u_hat[0] + g.ky * u_hat[1] + g.kz * u_hat[2]
    assert np.max(np.abs(div)) < 1e-10
    assert np.allclose(O.project(u_hat, g), u_hat)


# --- RHS / RK4: reference tests/test_navierstokes.py + brute force ---------


def test_beltrami_rhs_is_pure_viscous_decay(grid32):
    u0 = O.beltrami(grid32)
    du = O.make_ns_rhs(grid32)(u0, O.NSParams(nu=NU), 0.0)
    assert np.allclose(du, -NU * u0, atol=1e-10)


def test_rhs_nonlinear_term_brute_force():
    """At 8^3 the rotational term P[dealias(F(u x w))] is recomputed with
    direct physical-space sums (no FFT library) and compared."""
    Nn = 8
    g = O.make_grid(Nn)
    u_hat = O.project(dealiased_random_spectral(5, g), g)
    du = O.make_ns_rhs(g)(u_hat, O.NSParams(nu=0.0), 0.0)
    # full Hermitian spectrum on the 8^3 box
    full = np.zeros((3, Nn, Nn, Nn), complex)
    full[..., : Nn // 2 + 1] = u_hat
    for k in range(1, Nn // 2):
        full[..., Nn - k] = np.conj(u_hat[:, (-np.arange(Nn)) % Nn][:, :, (-np.arange(Nn)) % Nn][..., k])
    n = np.array([i if i < Nn // 2 else i - Nn for i in range(Nn)], float)
    kx, ky, kz = n[:, None, None], n[None, :, None], n[None, None, :]
    wf = 1j * np.stack((ky * full[2] - kz * full[1], kz * full[0] - kx * full[2], kx * full[1] - ky * full[0]))
    xs = np.arange(Nn) * 2 * np.pi / Nn
    E = np.exp(1j * np.outer(xs, n))                      # e^{i x k}
    def phys(a):
        return np.einsum("xi,yj,zk,cijk->cxyz", E, E, E, a).real / Nn**3
    u, w = phys(full), phys(wf)
    nl = np.stack((u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]))
    Ec = np.conj(E)
    nl_hat = np.einsum("xi,yj,zk,cxyz->cijk", Ec, Ec, Ec, nl)[..., : Nn // 2 + 1]
    ref = O.project(O.dealias(nl_hat, g), g)
    assert np.allclose(du, ref, atol=1e-11)


def test_beltrami_decay_matches_analytic(grid32):
    u0 = O.beltrami(grid32)
    rhs = O.make_ns_rhs(grid32)
    dt, n = 1e-2, 100
    u = O.rollout(rhs, u0, O.NSParams(nu=NU), dt, n)
    exact = np.exp(-NU * dt * n) * u0
    assert np.linalg.norm(u - exact) / np.linalg.norm(exact) < 1e-10
    z = -NU * dt                                          # RK4 amplification polynomial
    amp = (1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24) ** n
    assert np.linalg.norm(u - amp * u0) / np.linalg.norm(u0) < 1e-13


def test_rollout_preserves_solver_invariants(grid32):
    g = grid32
    u0 = O.project(dealiased_random_spectral(3, g), g)
    u = O.rollout(O.make_ns_rhs(g), u0, O.NSParams(nu=0.02), 1e-3, 10)
    div = g.kx * u[0] + g.ky * u[1] + g.kz * u[2]
    assert np.max(np.abs(div)) < 1e-8
    assert np.allclose(u * g.dealias, u)


def test_inviscid_energy_conservation(grid32):
    g = grid32
    u0 = O.project(dealiased_random_spectral(4, g), g)
    u = O.rollout(O.make_ns_rhs(g), u0, O.NSParams(nu=0.0), 1e-3, 20)
    assert np.isclose(O.energy(u, g), O.energy(u0, g), rtol=1e-8)


# --- forcings + budget: reference tests/test_budget.py ----------------------


def _measured_power(u_hat, term, g):
    return float(np.sum(g.weights * np.sum(np.real(np.conj(u_hat) * term), axis=0))) / O._ntot2(g)


def test_injection_exact_per_forcing(grid32):
    g = grid32
    u_hat = O.random_field(0, g)
    assert float(O.NoForcing().injection(u_hat, g)) == 0.0
    f = O.BandForcing(eps=0.3, kmin=1.0, kmax=3.0)
    assert float(f.injection(u_hat, g)) == 0.3
    assert float(f.injection(np.zeros_like(u_hat), g)) == 0.0
    assert np.isclose(_measured_power(u_hat, f.term(u_hat, g, 0.0), g), 0.3)
    lf = O.LinearForcing(A=0.1333)
    assert np.isclose(lf.injection(u_hat, g), 2 * 0.1333 * O.energy(u_hat, g))
    assert np.isclose(_measured_power(u_hat, lf.term(u_hat, g, 0.0), g), float(lf.injection(u_hat, g)))
    cf = O.LundgrenForcing(nu=NU, energy=0.5)
    assert np.isclose(cf.injection(u_hat, g), O.dissipation(u_hat, g, NU))
    assert np.isclose(_measured_power(u_hat, cf.term(u_hat, g, 0.0), g), O.dissipation(u_hat, g, NU))
    assert np.abs(cf.term(np.zeros_like(u_hat), g, 0.0)).max() == 0.0


@pytest.mark.parametrize("kmin,kmax", [(2.0, 3.0), (1.0, 2.0), (np.sqrt(5.0), 3.0)])
def test_band_forcing_edges_closed_form(kmin, kmax):
    """forcing.py:135-136 forces kmin^2 <= k2 <= kmax^2, both edges included.
    One mode per integer k2 = 1..12 (all at kz > 0, Hermitian weight 2, on a
    16^3 unit box where k2 is exact): the term is c a on exactly the modes
    with kmin^2 <= k2 <= kmax^2 (computed here from the integer mode numbers),
    0 elsewhere, with c = eps / (2 E_band), 2 E_band = sum of 2 |a|^2 / Ntot^2
    over those modes. A strict inequality at either edge, or a mask on |k|
    instead of k2, moves a mode across and fails."""
    g = O.make_grid(16)
    u = np.zeros((3, 16, 16, 9), dtype=np.complex128)
    rng = np.random.default_rng(7)
    forced, amps = [], {}
    for k2 in range(1, 13):
        n = next(((ix, iy, iz) for iz in range(1, 5) for iy in range(0, 5) for ix in range(0, 5)
                  if ix * ix + iy * iy + iz * iz == k2), None)
        if n is None:  # k2 = 7 is not a sum of three squares
            continue
        a = rng.standard_normal(3) + 1j * rng.standard_normal(3)
        u[:, n[0], n[1], n[2]] = a
        amps[n] = a
        if kmin * kmin <= k2 <= kmax * kmax:
            forced.append(n)
    assert forced
    two_e = sum(2 * np.sum(np.abs(amps[n]) ** 2) for n in forced) / float(16**3) ** 2
    f = O.BandForcing(eps=0.3, kmin=kmin, kmax=kmax)
    term = f.term(u, g, 0.0)
    expect = np.zeros_like(u)
    for n in forced:
        expect[:, n[0], n[1], n[2]] = (0.3 / two_e) * amps[n]
    assert np.max(np.abs(term - expect)) <= 1e-12 * np.max(np.abs(expect))
    assert float(f.injection(u, g)) == 0.3


def test_band_forcing_guard_is_dtype_eps():
    """forcing.py:141-143: no forcing when 2 E_band <= finfo(dtype).eps."""
    g = O.make_grid(16)
    u = np.zeros((3, 16, 16, 9), dtype=np.complex128)
    f = O.BandForcing(eps=0.3, kmin=1.0, kmax=1.0)
    # one mode at k2 = 1, kz = 1 (weight 2): 2 E_band = 2 |a|^2 / Ntot^2
    u[0, 0, 0, 1] = np.sqrt(0.4 * np.finfo(np.float64).eps) * 16**3
    assert np.abs(f.term(u, g, 0.0)).max() == 0.0 and float(f.injection(u, g)) == 0.0
    u[0, 0, 0, 1] = np.sqrt(0.6 * np.finfo(np.float64).eps) * 16**3
    assert np.abs(f.term(u, g, 0.0)).max() > 0.0 and float(f.injection(u, g)) == 0.3


def test_lundgren_holds_energy(grid32):
    """dE/dt = 2AE - eps = 0 for the balance rule: E stays at its initial
    value up to RK4 truncation; with tau the energy relaxes toward E*."""
    g = grid32
    u0 = O.random_field(1, g, O.config_spectrum)
    E0 = O.energy(u0, g)
    rhs = O.make_ns_rhs(g)
    u = O.rollout(rhs, u0, O.NSParams(nu=NU, forcing=O.LundgrenForcing(nu=NU)), 2e-3, 20)
    assert abs(O.energy(u, g) - E0) < 1e-9
    u = O.rollout(rhs, u0, O.NSParams(nu=NU, forcing=O.LundgrenForcing(nu=NU, energy=0.6, tau=0.05)), 2e-3, 50)
    Et = 0.6 - (0.6 - E0) * np.exp(-0.1 / 0.05)
    assert abs(O.energy(u, g) - Et) < 1e-6


def test_budget_residual_small_over_decay(grid32):
    """reference tests/test_budget.py:82-95."""
    g = grid32
    u = O.random_field(2, g)
    rhs = O.make_ns_rhs(g)
    dt, n = 1e-3, 50
    E, eps = [O.energy(u, g)], [O.dissipation(u, g, NU)]
    for _ in range(n):
        u = O.rk4_step(rhs, u, O.NSParams(nu=NU), dt, 0.0)
        E.append(O.energy(u, g)); eps.append(O.dissipation(u, g, NU))
    E, eps = np.array(E), np.array(eps)
    resid = (E[2:] - E[:-2]) / (2 * dt) + eps[1:-1]
    assert np.max(np.abs(resid)) < 1e-4 * eps.max()


# --- diagnostics: reference tests/test_diagnostics.py ----------------------


def test_parseval(grid32):
    u = np.random.default_rng(0).standard_normal((3, N, N, N))
    assert np.isclose(O.energy(O.to_spectral(u, grid32), grid32), 0.5 * np.sum(u**2) / N**3)


def test_spectrum_sums_to_energy(grid32):
    u_hat = O.random_field(1, grid32)
    _, Es = O.energy_spectrum(u_hat, grid32)
    assert np.isclose(Es.sum(), O.energy(u_hat, grid32))
    ks, E1, E2, E3 = O.component_spectra(u_hat, grid32)
    assert np.allclose(E1 + E2 + E3, Es)


def test_random_field_matches_target_spectrum(grid32):
    u_hat = O.random_field(2, grid32)
    ks, Es = O.energy_spectrum(u_hat, grid32)
    target = O.default_spectrum(ks)
    assert np.allclose(Es[1: N // 3], target[1: N // 3], rtol=1e-8)


def test_config_spectrum_energy_is_half():
    g = O.make_grid(64)
    u_hat = O.random_field(0, g, O.config_spectrum)
    assert abs(O.energy(u_hat, g) - 0.5) < 1e-12
    div = g.kx * u_hat[0] + g.ky * u_hat[1] + g.kz * u_hat[2]
    scale = np.max(np.sqrt(g.k2) * np.sqrt(np.sum(np.abs(u_hat) ** 2, axis=0)))
    assert np.max(np.abs(div)) < 1e-14 * scale and np.array_equal(u_hat * g.dealias, u_hat)


def test_beltrami_scales(grid32):
    u_hat = O.beltrami(grid32)
    assert np.isclose(O.enstrophy(u_hat, grid32), O.energy(u_hat, grid32))
    assert np.isclose(O.dissipation(u_hat, grid32, NU), 2 * NU * O.energy(u_hat, grid32))
    assert np.isclose(np.sqrt(5 * O.energy(u_hat, grid32) / O.enstrophy(u_hat, grid32)), np.sqrt(5))


# --- initial-condition noise (spec delta D4) --------------------------------


_M0, _M1 = 0xD2E7470EE14C6C93, 0xCA5A826395121157
_W0, _W1 = 0x9E3779B97F4A7C15, 0xBB67AE8584CAA73B
_MASK = (1 << 64) - 1


def _philox4x64_10(ctr, key):
    """Philox4x64-10 from its published definition (Salmon et al., SC'11):
    S-box = two 64x64->128 multiplies, 10 rounds, Weyl key schedule."""
    c, k = list(ctr), list(key)
    for _ in range(10):
        p0, p1 = _M0 * c[0], _M1 * c[2]
        c = [(p1 >> 64) ^ c[1] ^ k[0], p1 & _MASK, (p0 >> 64) ^ c[3] ^ k[1], p0 & _MASK]
        k = [(k[0] + _W0) & _MASK, (k[1] + _W1) & _MASK]
    return c


def test_philox_stream_is_documented_algorithm():
    raw = np.random.Philox(key=[7, 0], counter=0).random_raw(12)
    mine = sum((_philox4x64_10([b + 1, 0, 0, 0], [7, 0]) for b in range(3)), [])
    assert [int(x) for x in raw] == mine


def test_philox_uniform_mapping():
    raw = np.random.Philox(key=[3, 0], counter=0).random_raw(8)
    u64 = O.philox_uniform(3, (8,), np.float64)
    u32 = O.philox_uniform(3, (8,), np.float32)
    for i in range(8):
        assert u64[i] == (int(raw[i]) >> 11) / 2.0**53 - 0.5
        assert u32[i] == np.float32((int(raw[i]) >> 40) / 2.0**24 - 0.5)
    assert u64.min() >= -0.5 and u64.max() < 0.5


ALL_FORCINGS = [O.NoForcing(), O.LinearForcing(A=0.15), O.BandForcing(eps=0.3, kmin=1.0, kmax=3.0),
                O.LundgrenForcing(nu=0.02), O.LundgrenForcing(nu=0.02, energy=0.6, tau=0.3)]


def _rand_c(seed, shape):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("fi", range(len(ALL_FORCINGS)))
@pytest.mark.parametrize("dealiased", [True, False])
def test_rhs_vjp_matches_central_differences(fi, dealiased):
    """<vjp(u, v), d> equals the central difference of <v, rhs(u)> along d
    (dJ = Re sum conj(g) du) for every forcing, dense and dealiased inputs."""
    g = O.make_grid(16)
    u = O.random_field(4, g, O.config_spectrum)
    if not dealiased:
        u = u + 1e-2 * np.abs(u).max() * _rand_c(5, u.shape)
    v, d = _rand_c(6, u.shape), _rand_c(7, u.shape)
    params = O.NSParams(nu=0.02, forcing=ALL_FORCINGS[fi])
    rhs, vjp = O.make_ns_rhs(g), O.make_ns_rhs_vjp(g)
    eps = 1e-6 * np.abs(u).max() / np.abs(d).max()
    fd = (O._inner(v, rhs(u + eps * d, params, 0.0)) - O._inner(v, rhs(u - eps * d, params, 0.0))) / (2 * eps)
    an = O._inner(vjp(u, v, params, 0.0), d)
    assert abs(an - fd) < 1e-7 * abs(fd)


def test_rk4_step_and_rollout_vjp_match_central_differences():
    g = O.make_grid(16)
    u = O.random_field(8, g, O.config_spectrum)
    params = O.NSParams(nu=0.02, forcing=O.LundgrenForcing(nu=0.02))
    rhs, vjp = O.make_ns_rhs(g), O.make_ns_rhs_vjp(g)
    ub, d = _rand_c(9, u.shape), _rand_c(10, u.shape)
    eps = 1e-6 * np.abs(u).max() / np.abs(d).max()
    J = lambda x: O._inner(ub, O.rk4_step(rhs, x, params, 2e-3, 0.0))  # noqa: E731
    fd = (J(u + eps * d) - J(u - eps * d)) / (2 * eps)
    an = O._inner(O.rk4_step_vjp(rhs, vjp, u, ub, params, 2e-3, 0.0), d)
    assert abs(an - fd) < 1e-7 * abs(fd)
    Jn = lambda x: O._inner(ub, O.rollout(rhs, x, params, 2e-3, 7))  # noqa: E731
    fd = (Jn(u + eps * d) - Jn(u - eps * d)) / (2 * eps)
    for ck in (1, 3, 20):
        an = O._inner(O.rollout_vjp(rhs, vjp, u, ub, params, 2e-3, 7, checkpoint=ck), d)
        assert abs(an - fd) < 1e-7 * abs(fd)


PARAM_CASES = [
    ("nu", 0, lambda x: O.NSParams(nu=x, forcing=O.LinearForcing(A=0.15)), 0.02),
    ("A", 1, lambda x: O.NSParams(nu=0.02, forcing=O.LinearForcing(A=x)), 0.15),
    ("nu", 0, lambda x: O.NSParams(nu=x, forcing=O.BandForcing(eps=0.3, kmin=1.0, kmax=3.0)), 0.02),
    ("eps", 2, lambda x: O.NSParams(nu=0.02, forcing=O.BandForcing(eps=x, kmin=1.0, kmax=3.0)), 0.3),
    ("nu", 0, lambda x: O.NSParams(nu=x), 0.02),
]


@pytest.mark.parametrize("case", range(len(PARAM_CASES)))
def test_parameter_gradients_match_central_differences(case):
    """d/dnu, d/dA, d/deps of J = Re<ub, rollout(u, 5 steps)> (and of one RHS)
    from the hand adjoint against central differences in the parameter (the
    gradients the reference takes with jax.grad through NSParams,
    tests/test_grad.py:43-45, 59-74). The other slots stay 0."""
    name, slot, make, x0 = PARAM_CASES[case]
    g = O.make_grid(16)
    u = O.random_field(11, g, O.config_spectrum)
    ub = _rand_c(12, u.shape)
    rhs, vjp = O.make_ns_rhs(g), O.make_ns_rhs_vjp(g)
    h = 1e-4 * x0             # central differences: O(h^2) truncation, round-off well below
    # one RHS, also on a state with content outside the 2/3 box and a
    # divergent part (where the P M in the forcing derivatives matters)
    rough = u + 1e-2 * np.abs(u).max() * _rand_c(13, u.shape)
    for x in (u, rough):
        J1 = lambda p: O._inner(ub, rhs(x, make(p), 0.0))  # noqa: E731
        fd = (J1(x0 + h) - J1(x0 - h)) / (2 * h)
        pg = O.rhs_param_vjp(x, ub, make(x0), g)
        assert abs(pg[slot] - fd) < 1e-7 * abs(fd)
    f = make(x0).forcing
    live = {0} | ({1} if isinstance(f, O.LinearForcing) else set()) | ({2} if isinstance(f, O.BandForcing) else set())
    assert set(np.flatnonzero(pg)) == live
    Jn = lambda x: O._inner(ub, O.rollout(rhs, u, make(x), 2e-3, 5))  # noqa: E731
    fd = (Jn(x0 + h) - Jn(x0 - h)) / (2 * h)
    for ck in (1, 2):
        pg = np.zeros(3)
        O.rollout_vjp(rhs, vjp, u, ub, make(x0), 2e-3, 5, checkpoint=ck, pgrad=pg)
        assert abs(pg[slot] - fd) < 1e-7 * abs(fd), (name, pg, fd)


def test_taylor_green_closed_form():
    """fields.py:71-87: the field is band-limited (|n_i| = 1), so its physical
    values come back exactly and it is divergence-free."""
    g = O.make_grid(16)
    u = O.to_physical(O.taylor_green(g), g)
    x = (np.arange(16) * 2 * np.pi / 16).reshape(-1, 1, 1)
    y, z = x.reshape(1, -1, 1), x.reshape(1, 1, -1)
    assert np.allclose(u[0], np.sin(x) * np.cos(y) * np.cos(z), atol=1e-14)
    assert np.allclose(u[1], -np.cos(x) * np.sin(y) * np.cos(z), atol=1e-14)
    assert np.allclose(u[2], 0, atol=1e-14)
    uh = O.taylor_green(g)
    assert np.max(np.abs(g.kx * uh[0] + g.ky * uh[1] + g.kz * uh[2])) < 1e-10


def test_forced_256_record_matches_the_oracle():
    """tests/golden/forced_256_lundgren.npz (written by make_forced_256.py from
    this oracle) is current: its step-0 spectrum, energy and sampled modes
    are the oracle's random_field(0) at 256^3, E = 0.5 from the C k^4 exp(-2
    (k/4)^2) normalisation, eps = 2 nu Omega, and E stays at 0.5 to round-off
    under the constant-energy forcing over the 10 recorded steps."""
    import os
    rec = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                               "forced_256_lundgren.npz"))
    g = O.make_grid(256)
    u0 = O.random_field(int(rec["seed"]), g, O.config_spectrum)
    assert np.array_equal(u0.ravel()[rec["index"]], rec["sample0"])
    _, E = O.energy_spectrum(u0, g)
    assert np.allclose(E, rec["spectra"][0][1:], rtol=1e-13, atol=0)
    assert abs(rec["E"][0] - 0.5) < 1e-14
    assert np.allclose(rec["eps"], 2 * float(rec["nu"]) * rec["Omega"], rtol=1e-15)
    assert np.max(np.abs(rec["E"] - 0.5)) < 1e-12
    assert np.allclose(rec["spectra"].sum(axis=1), rec["E"], rtol=1e-12)


# --- integrating-factor (Lawson) RK4: the scheme navierstokes.py:100-103 leaves room for ---


def test_if_rk4_reduces_to_rk4_without_viscosity():
    g = O.make_grid(16)
    u = O.random_field(3, g, O.config_spectrum)
    rhs, p = O.make_ns_rhs(g), O.NSParams(nu=0.0, forcing=O.LinearForcing(A=0.1))
    assert np.allclose(O.if_rk4_step(rhs, u, p, 1e-2, 0.0), O.rk4_step(rhs, u, p, 1e-2, 0.0),
                       rtol=0, atol=1e-14 * np.abs(u).max())


def test_if_rk4_is_exact_for_pure_viscous_decay():
    """Beltrami (u x w = 0): the stage derivative N vanishes, so the scheme
    reproduces exp(-nu t) at any step size, where RK4 carries its
    truncation error (reference tests/test_navierstokes.py:22-32)."""
    g = O.make_grid(16)
    u0 = O.beltrami(g)
    rhs, p = O.make_ns_rhs(g), O.NSParams(nu=0.5)
    u = O.rollout(rhs, u0, p, 0.4, 5, step=O.if_rk4_step)
    exact = np.exp(-0.5 * 2.0) * u0
    assert np.max(np.abs(u - exact)) < 1e-13 * np.max(np.abs(u0))
    u_rk4 = O.rollout(rhs, u0, p, 0.4, 5)
    assert np.max(np.abs(u_rk4 - exact)) > 1e-6 * np.max(np.abs(u0))


def test_if_rk4_is_fourth_order():
    """Global error at t = 0.2 against a 64-step solution of the same
    scheme: halving dt divides it by about 16."""
    g = O.make_grid(16)
    u0 = O.random_field(5, g, O.config_spectrum)
    rhs, p = O.make_ns_rhs(g), O.NSParams(nu=0.05, forcing=O.LundgrenForcing(nu=0.05))
    ref = O.rollout(rhs, u0, p, 0.2 / 64, 64, step=O.if_rk4_step)
    e = [np.linalg.norm(O.rollout(rhs, u0, p, 0.2 / n, n, step=O.if_rk4_step) - ref) for n in (4, 8)]
    assert 12 < e[0] / e[1] < 20


def test_if_rk4_is_stable_where_rk4_is_not():
    """nu k2 dt = 3.75 on the highest retained modes (beyond RK4's real-axis
    limit 2.79): the integrating factor keeps the viscous decay, RK4 grows."""
    g = O.make_grid(16)
    u0 = O.random_field(6, g, O.config_spectrum)
    rhs, p = O.make_ns_rhs(g), O.NSParams(nu=1.0)
    E0 = O.energy(u0, g)
    assert O.energy(O.rollout(rhs, u0, p, 0.05, 10, step=O.if_rk4_step), g) < E0
    with np.errstate(all="ignore"):
        E_rk4 = O.energy(O.rollout(rhs, u0, p, 0.05, 10), g)
    assert not E_rk4 <= 10 * E0          # grows without bound (overflow to nan included)
