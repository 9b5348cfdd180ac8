"""Parity of the HIP path (called through the C ABI via the hit3d package)
with the CPU oracle on the same inputs.

Bars: tables / masks / shell indices / noise bit-exact; float64 kernels at
round-off (<= 1e-11 relative); float32 kernels at the tolerance written in
each test (state after 100 steps <= 1e-5 relative L2 against float64)."""

import numpy as np
import pytest
import torch

import hit3d_oracle as O

pytestmark = pytest.mark.gpu

h3 = pytest.importorskip("hit3d")

DT = {np.float32: torch.float32, np.float64: torch.float64}


def _grids(shape, dtype):
    return O.make_grid(*shape, dtype=dtype), h3.make_grid(*shape, dtype=DT[dtype])


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = b.detach().cpu().numpy() if isinstance(b, torch.Tensor) else np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(np.asarray(b).ravel()), 1e-300))


def dev(x, grid):
    return torch.from_numpy(np.ascontiguousarray(x)).to(grid.device)


def random_spectral(seed, og, dealiased=True, ncomp=3):
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((ncomp, og.Nx, og.Ny, og.Nz)).astype(og.dtype)
    uh = O.to_spectral(u, og)
    return O.dealias(uh, og) if dealiased else uh


# --- tables: bit-exact ----------------------------------------------------------


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(16, 16, 16), (64, 64, 64), (16, 8, 4), (32, 16, 64), (256, 256, 256)])
def test_tables_bit_exact(shape, dtype):
    og, g = _grids(shape, dtype)
    assert np.array_equal(g.k2.cpu().numpy(), og.k2)
    assert np.array_equal(g.inv_k2.cpu().numpy(), og.inv_k2)
    assert np.array_equal(g.dealias.cpu().numpy(), og.dealias)
    shell, _, _ = O.shells(og)
    assert np.array_equal(g.shell.cpu().numpy(), shell)
    assert np.array_equal(g.kx.cpu().numpy(), og.kx) and np.array_equal(g.kz.cpu().numpy(), og.kz)
    assert np.array_equal(g.weights.cpu().numpy(), og.weights)


def test_tables_bit_exact_non_unit_box():
    og = O.make_grid(32, 16, 8, Lx=3.0, Ly=1.7, Lz=5.5, dtype=np.float32)
    g = h3.make_grid(32, 16, 8, Lx=3.0, Ly=1.7, Lz=5.5, dtype=torch.float32)
    assert np.array_equal(g.k2.cpu().numpy(), og.k2)
    assert np.array_equal(g.inv_k2.cpu().numpy(), og.inv_k2)
    assert np.array_equal(g.shell.cpu().numpy(), O.shells(og)[0])


# --- noise: bit-exact Philox ----------------------------------------------------------


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_uniform_noise_bit_exact(dtype):
    og, g = _grids((16, 8, 32), dtype)
    ref = O.philox_uniform(123, (3, 16, 8, 32), dtype)
    assert np.array_equal(h3.uniform_noise(123, g).cpu().numpy(), ref)


# --- transforms -----------------------------------------------------------------------


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 2e-6)])
@pytest.mark.parametrize("shape", [(4, 4, 4), (8, 8, 8), (16, 8, 4), (32, 64, 16), (64, 64, 64),
                                   (128, 128, 128), (256, 32, 512), (1024, 8, 8), (8, 1024, 4),
                                   (4, 8, 1024)])
def test_transforms_match_numpy(shape, dtype, tol):
    og, g = _grids(shape, dtype)
    u = np.random.default_rng(0).standard_normal((3,) + shape).astype(dtype)
    uh = h3.to_spectral(dev(u, g), g)
    assert rel(uh, O.to_spectral(u, og)) < tol
    # irfftn of a generic (non-Hermitian) half spectrum follows c2r semantics
    rng = np.random.default_rng(1)
    s = (2,) + shape[:2] + (shape[2] // 2 + 1,)
    vh = (rng.standard_normal(s) + 1j * rng.standard_normal(s)).astype(O.complex_dtype(og))
    assert rel(h3.to_physical(dev(vh, g), g), O.to_physical(vh, og)) < tol
    assert rel(h3.to_physical(uh, g), u) < tol


def test_transform_single_component_and_odd_count():
    og, g = _grids((16, 16, 16), np.float64)
    u = np.random.default_rng(3).standard_normal((16, 16, 16))
    assert rel(h3.to_spectral(dev(u, g), g), O.to_spectral(u[None], og)[0]) < 1e-13
    u5 = np.random.default_rng(4).standard_normal((5, 16, 16, 16))
    uh5 = h3.to_spectral(dev(u5, g), g)
    assert rel(uh5, O.to_spectral(u5, og)) < 1e-13
    assert rel(h3.to_physical(uh5, g), u5) < 1e-13


# --- elementwise operators ----------------------------------------------------------------


def test_dealias_project_derivatives():
    og, g = _grids((32, 16, 32), np.float64)
    uh = random_spectral(5, og, dealiased=False)
    d = dev(uh, g)
    assert rel(h3.dealias(d, g), O.dealias(uh, og)) == 0.0
    assert rel(h3.project(d, g), O.project(uh, og)) < 1e-15
    for f, of in ((h3.ddx, O.ddx), (h3.ddy, O.ddy), (h3.ddz, O.ddz), (h3.laplacian, O.laplacian)):
        assert rel(f(d, g), of(uh, og)) < 1e-15


# --- RHS ---------------------------------------------------------------------------------------

FORCINGS = [
    (lambda: O.NoForcing(), lambda: h3.NoForcing()),
    (lambda: O.LinearForcing(A=0.15), lambda: h3.LinearForcing(A=0.15)),
    (lambda: O.BandForcing(eps=0.3, kmin=1.0, kmax=3.0), lambda: h3.BandForcing(eps=0.3, kmin=1.0, kmax=3.0)),
    (lambda: O.LundgrenForcing(nu=0.02, energy=0.5, tau=0.3), lambda: h3.LundgrenForcing(nu=0.02, energy=0.5, tau=0.3)),
]


@pytest.mark.parametrize("fi", [1, 2, 3])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-14), (np.float32, 1e-6)])
def test_forcing_term_matches_oracle(fi, dtype, tol):
    """forcing.term (hit3d_forcing_term: coefficient reduced in FP64 on the
    device) against the oracle's term (forcing.py:74-163), on a state with
    content outside the 2/3 box and on an anisotropic box; BandForcing's
    zero-energy guard gives an exact zero field."""
    for shape, box in (((32, 32, 32), (2 * np.pi,) * 3), ((16, 32, 8), (3.0, 1.7, 5.5))):
        og = O.make_grid(*shape, Lx=box[0], Ly=box[1], Lz=box[2], dtype=dtype)
        g = h3.make_grid(*shape, Lx=box[0], Ly=box[1], Lz=box[2], dtype=DT[dtype])
        uh = random_spectral(30 + fi, og, dealiased=False)
        of, gf = FORCINGS[fi][0](), FORCINGS[fi][1]()
        got = gf.term(dev(uh, g), g, 0.0)
        assert rel(got, of.term(uh, og, 0.0)) < tol
    if fi == 2:   # no energy in the band: the guard zeroes the term exactly
        og, g = _grids((16, 16, 16), dtype)
        uh = random_spectral(40, og)
        uh = np.where((og.k2 >= 1.0) & (og.k2 <= 9.0), 0, uh)
        assert float(gf.term(dev(uh, g), g, 0.0).abs().max()) == 0.0
        assert np.abs(of.term(uh, og, 0.0)).max() == 0.0


@pytest.mark.parametrize("fi", range(len(FORCINGS)))
@pytest.mark.parametrize("dealiased", [True, False])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 3e-6)])
def test_rhs_matches_oracle(fi, dealiased, dtype, tol):
    og, g = _grids((32, 32, 32), dtype)
    uh = O.project(random_spectral(7 + fi, og, dealiased), og)
    of, gf = FORCINGS[fi]
    ref = O.make_ns_rhs(og)(uh.astype(np.complex128) if dtype == np.float64 else uh,
                            O.NSParams(nu=0.02, forcing=of()), 0.0)
    du = h3.make_ns_rhs(g)(dev(uh, g), h3.NSParams(nu=0.02, forcing=gf()), 0.0)
    assert h3.is_dealiased(dev(uh, g), g) == dealiased
    assert rel(du, ref) < tol


@pytest.mark.parametrize("shape", [(16, 8, 4), (8, 32, 16), (64, 16, 32)])
def test_rhs_anisotropic(shape):
    og, g = _grids(shape, np.float64)
    uh = O.project(random_spectral(11, og), og)
    ref = O.make_ns_rhs(og)(uh, O.NSParams(nu=0.01), 0.0)
    assert rel(h3.make_ns_rhs(g)(dev(uh, g), h3.NSParams(nu=0.01), 0.0), ref) < 1e-12


# The fused pruned passes at the large transform sizes (256-, 512- and
# 1024-point kernels, padded kz rows) along each axis in turn, on grids small
# enough for the oracle.
LARGE_AXIS = [(256, 8, 8), (8, 256, 8), (8, 8, 256), (512, 8, 4), (8, 512, 4), (4, 8, 512),
              (1024, 8, 4), (4, 1024, 8), (8, 4, 1024)]


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 5e-6)])
@pytest.mark.parametrize("shape", LARGE_AXIS)
def test_rhs_large_axis_pruned(shape, dtype, tol):
    og, g = _grids(shape, dtype)
    uh = O.project(random_spectral(12, og), og)
    params = (O.NSParams(nu=0.01, forcing=O.LundgrenForcing(nu=0.01, energy=0.5, tau=0.3)),
              h3.NSParams(nu=0.01, forcing=h3.LundgrenForcing(nu=0.01, energy=0.5, tau=0.3)))
    ref = O.make_ns_rhs(og)(uh, params[0], 0.0)
    assert rel(h3.make_ns_rhs(g)(dev(uh, g), params[1], 0.0), ref) < tol


@pytest.mark.parametrize("dtype,tol", [(np.float32, 5e-6), (np.float64, 1e-12)])
@pytest.mark.parametrize("shape", [(32, 64, 64), (64, 32, 64), (16, 64, 64), (256, 16, 64)])
def test_rhs_and_rollout_pencil_layouts(shape, dtype, tol):
    """The forward RHS's pencil buffer is y-major when its kz rows are not
    whole 128-byte lines (float32 at Nz = 64: 24 modes, 192 bytes) and
    x-major otherwise (float64: 384 bytes); both against the oracle on boxes
    with Nx != Ny, where an x/y mix-up of the y-major addressing (row stride,
    pencil groups along x, padded rows) would show, for the RHS and a
    3-step forced rollout."""
    og, g = _grids(shape, dtype)
    uh = O.project(random_spectral(14, og), og)
    op = O.NSParams(nu=0.01, forcing=O.LundgrenForcing(nu=0.01, energy=0.5, tau=0.3))
    hp = h3.NSParams(nu=0.01, forcing=h3.LundgrenForcing(nu=0.01, energy=0.5, tau=0.3))
    ref = O.make_ns_rhs(og)(uh, op, 0.0)
    assert rel(h3.make_ns_rhs(g)(dev(uh, g), hp, 0.0), ref) < tol
    ref = O.rollout(O.make_ns_rhs(og), uh, op, 1e-3, 3)
    got = h3.rollout(h3.make_ns_rhs(g), dev(uh, g), hp, 1e-3, 3)
    assert rel(got, ref) < (2e-5 if dtype == np.float32 else 1e-12)


def test_rhs_and_rollout_in_two_x_halves_with_x_major_pencils():
    """A float32 pencil buffer between 192 and 384 MB makes the forward RHS
    run its y-inverse, z and y-forward passes over two halves of the x range
    (256^3 covers the y-major form); at (256, 512, 128) the kz rows (48 modes,
    384 bytes) are whole lines, so the halves go through the x-major buffer
    (302 MB). Against the float64 path on the same grid, which runs whole (its
    pencil buffer is 604 MB)."""
    shape = (256, 512, 128)
    og = O.make_grid(*shape, dtype=np.float64)
    uh = O.project(random_spectral(15, og), og)
    g32, g64 = h3.make_grid(*shape, dtype=torch.float32), h3.make_grid(*shape, dtype=torch.float64)
    hp = h3.NSParams(nu=0.01, forcing=h3.LundgrenForcing(nu=0.01, energy=0.5, tau=0.3))
    u32 = dev(uh.astype(np.complex64), g32)
    r32 = h3.make_ns_rhs(g32)(u32, hp, 0.0)
    r64 = h3.make_ns_rhs(g64)(dev(uh, g64), hp, 0.0)
    assert rel(r32, r64.cpu().numpy()) < 5e-6
    u32 = h3.rollout(h3.make_ns_rhs(g32), u32, hp, 1e-3, 2)
    u64 = h3.rollout(h3.make_ns_rhs(g64), dev(uh, g64), hp, 1e-3, 2)
    assert rel(u32, u64.cpu().numpy()) < 2e-6


def test_rhs_and_rollout_in_sixteen_x_chunks_at_512():
    """The float32 512^3 pencil buffer (2.2 GB) makes the forward RHS run its
    y-inverse, z and y-forward passes over 16 chunks of 32 x-planes; against
    the float64 path (4.4 GB, run whole) from the same seeded field."""
    g32, g64 = h3.make_grid(512, dtype=torch.float32), h3.make_grid(512, dtype=torch.float64)
    u64 = h3.random_field(5, g64, h3.config_spectrum)
    u32 = u64.to(torch.complex64)
    hp = h3.NSParams(nu=1 / 2000, forcing=h3.LundgrenForcing(nu=1 / 2000, energy=0.5))
    r32 = h3.make_ns_rhs(g32)(u32, hp, 0.0)
    r64 = h3.make_ns_rhs(g64)(u64, hp, 0.0)
    err = float(torch.linalg.vector_norm((r32.to(torch.complex128) - r64).reshape(-1))
                / torch.linalg.vector_norm(r64.reshape(-1)))
    assert err < 5e-6, err
    del r32, r64
    u32 = h3.rollout(h3.make_ns_rhs(g32), u32, hp, 2e-3, 2)
    u64 = h3.rollout(h3.make_ns_rhs(g64), u64, hp, 2e-3, 2)
    err = float(torch.linalg.vector_norm((u32.to(torch.complex128) - u64).reshape(-1))
                / torch.linalg.vector_norm(u64.reshape(-1)))
    assert err < 2e-6, err


@pytest.mark.parametrize("shape", [(1024, 8, 4), (8, 8, 256), (4, 8, 512), (8, 512, 4)])
def test_rollout_large_axis_pruned(shape):
    """RK4 stage kernels and the per-stage FP64 coefficient at large N."""
    og, g = _grids(shape, np.float64)
    uh = O.project(random_spectral(13, og), og)
    op = O.NSParams(nu=0.01, forcing=O.LundgrenForcing(nu=0.01, energy=0.5))
    hp = h3.NSParams(nu=0.01, forcing=h3.LundgrenForcing(nu=0.01, energy=0.5))
    ref = O.rollout(O.make_ns_rhs(og), uh, op, 1e-3, 3)
    got = h3.rollout(h3.make_ns_rhs(g), dev(uh, g), hp, 1e-3, 3)
    assert rel(got, ref) < 1e-12


def test_beltrami_rhs_is_pure_viscous_decay():
    """reference tests/test_navierstokes.py:10-19 on the GPU (float64, dense path)."""
    og, g = _grids((32, 32, 32), np.float64)
    u0 = O.beltrami(og)
    du = h3.make_ns_rhs(g)(dev(u0, g), h3.NSParams(nu=0.05), 0.0)
    assert np.allclose(du.cpu().numpy(), -0.05 * u0, atol=1e-10)


# --- RK4 ------------------------------------------------------------------------------------------


def test_beltrami_decay_matches_analytic():
    """reference tests/test_navierstokes.py:22-32 (float64)."""
    og, g = _grids((32, 32, 32), np.float64)
    u0 = O.beltrami(og)
    u = h3.rollout(h3.make_ns_rhs(g), dev(u0, g), h3.NSParams(nu=0.05), 1e-2, 100)
    assert rel(u, np.exp(-0.05 * 1.0) * u0) < 1e-10


@pytest.mark.parametrize("fi", range(len(FORCINGS)))
def test_rollout_float64_matches_oracle(fi):
    og, g = _grids((32, 32, 32), np.float64)
    u0 = O.random_field(3, og, O.config_spectrum)
    of, gf = FORCINGS[fi]
    ref = O.rollout(O.make_ns_rhs(og), u0, O.NSParams(nu=0.02, forcing=of()), 2e-3, 10)
    u = h3.rollout(h3.make_ns_rhs(g), dev(u0, g), h3.NSParams(nu=0.02, forcing=gf()), 2e-3, 10)
    assert rel(u, ref) < 1e-11


def test_rk4_step_dense_path_matches_oracle():
    og, g = _grids((16, 16, 16), np.float64)
    u0 = O.project(random_spectral(21, og, dealiased=False), og)
    ref = O.rk4_step(O.make_ns_rhs(og), u0, O.NSParams(nu=0.03), 1e-3, 0.0)
    u = h3.rk4_step(h3.make_ns_rhs(g), dev(u0, g), h3.NSParams(nu=0.03), 1e-3, 0.0)
    assert rel(u, ref) < 1e-12


def test_decaying_64_float32_100_steps_within_1e5():
    """The decaying 64^3 case (nu = 1/400, dt = 2e-3, 100 steps): the float32
    GPU state against the float64 oracle from the same initial field."""
    og, g = _grids((64, 64, 64), np.float64)
    g32 = h3.make_grid(64, dtype=torch.float32)
    u0 = h3.random_field(0, g32, h3.config_spectrum)
    u0_64 = u0.cpu().numpy().astype(np.complex128)
    ref = O.rollout(O.make_ns_rhs(og), u0_64, O.NSParams(nu=1 / 400), 2e-3, 100)
    u = h3.rollout(h3.make_ns_rhs(g32), u0, h3.NSParams(nu=1 / 400), 2e-3, 100)
    err = rel(u.to(torch.complex128), ref)
    print(f"decaying 64^3 fp32 vs fp64 oracle after 100 steps: rel L2 = {err:.3e}")
    assert err < 1e-5


def test_rollout_invariants_and_inviscid_energy():
    """reference tests/test_navierstokes.py:35-54 (float64)."""
    og, g = _grids((32, 32, 32), np.float64)
    u0 = O.project(random_spectral(4, og), og)
    rhs = h3.make_ns_rhs(g)
    u = h3.rollout(rhs, dev(u0, g), h3.NSParams(nu=0.02), 1e-3, 10)
    un = u.cpu().numpy()
    div = og.kx * un[0] + og.ky * un[1] + og.kz * un[2]
    assert np.max(np.abs(div)) < 1e-8
    assert np.array_equal(un * og.dealias, un)
    u = h3.rollout(rhs, dev(u0, g), h3.NSParams(nu=0.0), 1e-3, 20)
    assert np.isclose(h3.energy(u, g), O.energy(u0, og), rtol=1e-8)


def test_lundgren_holds_energy_float32():
    g = h3.make_grid(64, dtype=torch.float32)
    u0 = h3.random_field(0, g, h3.config_spectrum)
    assert abs(h3.energy(u0, g) - 0.5) < 1e-6
    u = h3.rollout(h3.make_ns_rhs(g), u0, h3.NSParams(nu=1e-3, forcing=h3.LundgrenForcing(nu=1e-3)), 1e-3, 50)
    assert abs(h3.energy(u, g) - 0.5) < 2e-5


def test_rollout_is_pure_and_zero_steps():
    og, g = _grids((16, 16, 16), np.float64)
    u0 = dev(O.random_field(1, og), g)
    keep = u0.clone()
    h3.rollout(h3.make_ns_rhs(g), u0, h3.NSParams(nu=0.02), 1e-3, 3)
    assert torch.equal(u0, keep)
    assert torch.equal(h3.rollout(h3.make_ns_rhs(g), u0, h3.NSParams(nu=0.02), 1e-3, 0), keep)


# --- diagnostics ------------------------------------------------------------------------------------


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 1e-6)])
def test_diagnostics_match_oracle(dtype, tol):
    og, g = _grids((64, 32, 32), dtype)
    uh = random_spectral(9, og, dealiased=False)
    d = dev(uh, g)
    uh64 = uh.astype(np.complex128)
    og64 = O.make_grid(64, 32, 32)
    assert abs(h3.energy(d, g) / O.energy(uh64, og64) - 1) < tol
    assert abs(h3.enstrophy(d, g) / O.enstrophy(uh64, og64) - 1) < tol
    assert abs(h3.dissipation(d, g, 0.01) / O.dissipation(uh64, og64, 0.01) - 1) < tol
    ks, Es = h3.energy_spectrum(d, g)
    oks, oEs = O.energy_spectrum(uh64, og64)
    assert np.array_equal(ks, oks)
    assert np.max(np.abs(Es - oEs)) < tol * np.max(oEs)
    _, E1, E2, E3 = h3.component_spectra(d, g)
    _, oE1, oE2, oE3 = O.component_spectra(uh64, og64)
    for a, b in ((E1, oE1), (E2, oE2), (E3, oE3)):
        assert np.max(np.abs(a - b)) < tol * np.max(b)


def test_spectrum_deterministic_and_sums_to_energy():
    g = h3.make_grid(128, dtype=torch.float32)
    u = h3.random_field(5, g, h3.config_spectrum)
    _, E1 = h3.energy_spectrum(u, g)
    _, E2 = h3.energy_spectrum(u, g)
    assert np.array_equal(E1, E2)
    assert abs(E1.sum() - h3.energy(u, g)) < 1e-12


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 1e-6)])
def test_random_field_matches_oracle(dtype, tol):
    og, g = _grids((32, 32, 32), dtype)
    ref = O.random_field(2, og)
    u = h3.random_field(2, g)
    assert rel(u, ref) < tol
    ks, Es = h3.energy_spectrum(u, g)
    target = O.default_spectrum(ks)
    assert np.allclose(Es[1:10], target[1:10], rtol=max(tol, 1e-8) * 10)


def test_injection_and_budget_match_oracle():
    og, g = _grids((32, 32, 32), np.float64)
    uh = O.random_field(0, og)
    d = dev(uh, g)
    for of, gf in FORCINGS:
        assert np.isclose(gf().injection(d, g), float(of().injection(uh, og)), rtol=1e-12, atol=0)
    b = h3.energy_budget(d, g, 0.05, h3.BandForcing(eps=0.3, kmin=1.0, kmax=3.0))
    assert b["P"] == 0.3 and np.isclose(b["E"], O.energy(uh, og), rtol=1e-13)
    assert h3.BandForcing(eps=0.3, kmin=1.0, kmax=3.0).injection(torch.zeros_like(d), g) == 0.0


# --- error behaviour -------------------------------------------------------------------------------


def test_bad_inputs_raise():
    g = h3.make_grid(16, dtype=torch.float32)
    rhs = h3.make_ns_rhs(g)
    with pytest.raises(ValueError):
        rhs(torch.zeros((3, 16, 16, 8), dtype=torch.complex64, device="cuda"), h3.NSParams(nu=0.1), 0.0)
    with pytest.raises(ValueError):
        rhs(torch.zeros((3, 16, 16, 9), dtype=torch.complex128, device="cuda"), h3.NSParams(nu=0.1), 0.0)
    with pytest.raises(ValueError):
        h3.make_grid(12)
    with pytest.raises(ValueError):
        h3.rollout(rhs, torch.zeros((3, 16, 16, 9), dtype=torch.complex64, device="cuda"),
                   h3.NSParams(nu=0.1), 1e-3, -1)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 1e-6)])
def test_taylor_green_matches_oracle(dtype, tol):
    og, g = _grids((16, 32, 16), dtype)
    assert rel(h3.taylor_green(g), O.taylor_green(og)) < tol


# --- boxes with L != 2 pi (k0 != 1) --------------------------------------------------------------
# The wavenumbers scale with 2 pi / L (grid.py:105-107) and the spectrum bins
# by k0 = 2 pi / Lx (diagnostics.py:76-82): RHS, rollout and spectra on a cubic
# box of side 3 and on an anisotropic box.

BOXES = [((32, 32, 32), (3.0, 3.0, 3.0)), ((32, 16, 8), (3.0, 1.7, 5.5)), ((16, 32, 64), (9.0, 9.0, 9.0))]


@pytest.mark.parametrize("box", range(len(BOXES)))
@pytest.mark.parametrize("fi", range(len(FORCINGS)))
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 3e-6)])
def test_rhs_non_unit_box(box, fi, dtype, tol):
    shape, (lx, ly, lz) = BOXES[box]
    og = O.make_grid(*shape, Lx=lx, Ly=ly, Lz=lz, dtype=dtype)
    g = h3.make_grid(*shape, Lx=lx, Ly=ly, Lz=lz, dtype=DT[dtype])
    uh = O.project(random_spectral(30 + fi, og), og)
    of, gf = FORCINGS[fi]
    ref = O.make_ns_rhs(og)(uh, O.NSParams(nu=0.02, forcing=of()), 0.0)
    for dealiased in (True, False):   # pruned and dense layouts on the same input
        du = h3.make_ns_rhs(g)(dev(uh, g), h3.NSParams(nu=0.02, forcing=gf()), 0.0, dealiased=dealiased)
        assert rel(du, ref) < tol


@pytest.mark.parametrize("box", range(len(BOXES)))
def test_rollout_and_spectra_non_unit_box(box):
    shape, (lx, ly, lz) = BOXES[box]
    og = O.make_grid(*shape, Lx=lx, Ly=ly, Lz=lz)
    g = h3.make_grid(*shape, Lx=lx, Ly=ly, Lz=lz, dtype=torch.float64)
    uh = O.project(random_spectral(40 + box, og), og)
    op = O.NSParams(nu=0.01, forcing=O.LundgrenForcing(nu=0.01, energy=0.5, tau=0.5))
    hp = h3.NSParams(nu=0.01, forcing=h3.LundgrenForcing(nu=0.01, energy=0.5, tau=0.5))
    ref = O.rollout(O.make_ns_rhs(og), uh, op, 2e-3, 10)
    got = h3.rollout(h3.make_ns_rhs(g), dev(uh, g), hp, 2e-3, 10)
    assert rel(got, ref) < 1e-11
    ks, Es = h3.energy_spectrum(got, g)
    oks, oEs = O.energy_spectrum(ref, og)
    assert np.array_equal(ks, oks)           # (1..smax) * 2 pi / Lx
    assert np.max(np.abs(Es - oEs)) < 1e-12 * np.max(oEs)
    for a, b in zip(h3.component_spectra(got, g)[1:], O.component_spectra(ref, og)[1:]):
        assert np.max(np.abs(a - b)) < 1e-12 * np.max(b)
    assert abs(h3.energy(got, g) / O.energy(ref, og) - 1) < 1e-12
    assert abs(h3.dissipation(got, g, 0.01) / O.dissipation(ref, og, 0.01) - 1) < 1e-12


# --- BandForcing band edges (forcing.py:135-136: kmin^2 <= k2 <= kmax^2) -------------------------


@pytest.mark.parametrize("dealiased", [True, False])
def test_band_forcing_edges_closed_form(dealiased):
    """Four divergence-free single modes on a 16^3 box, at k2 = kmin^2 - 1,
    kmin^2, kmax^2 and kmax^2 + 1 for kmin = 2, kmax = 3: the forcing part of
    the RHS, rhs(band) - rhs(none), must equal c u on the two edge modes and
    vanish on the two outside ones, with c = eps / (2 E_band) and 2 E_band
    = sum over the two forced modes of 2 |a|^2 / Ntot^2 (all at kz > 0,
    Hermitian weight 2)."""
    og, g = _grids((16, 16, 16), np.float64)
    u = np.zeros((3, 16, 16, 9), dtype=np.complex128)
    modes = {  # (ix, iy, iz): (vector perpendicular to k, amplitude, forced?)
        (1, 1, 1): ((1.0, -1.0, 0.0), 0.7 + 0.2j, False),   # k2 = 3 = kmin^2 - 1
        (0, 0, 2): ((1.0, 0.0, 0.0), 0.4 - 0.9j, True),     # k2 = 4 = kmin^2
        (0, 0, 3): ((0.0, 1.0, 0.0), -1.1 + 0.3j, True),    # k2 = 9 = kmax^2
        (1, 0, 3): ((0.0, 1.0, 0.0), 0.5 + 0.5j, False),    # k2 = 10 = kmax^2 + 1
    }
    for (ix, iy, iz), (vec, amp, _) in modes.items():
        for c in range(3):
            u[c, ix, iy, iz] = vec[c] * amp * 1000.0
    eps = 0.3
    two_e_band = sum(2 * np.sum(np.abs(u[:, ix, iy, iz]) ** 2)
                     for (ix, iy, iz), (_, _, forced) in modes.items() if forced) / float(16**3) ** 2
    coef = eps / two_e_band
    rhs = h3.make_ns_rhs(g)
    d = dev(u, g)
    band = rhs(d, h3.NSParams(nu=0.0, forcing=h3.BandForcing(eps=eps, kmin=2.0, kmax=3.0)), 0.0,
               dealiased=dealiased)
    none = rhs(d, h3.NSParams(nu=0.0), 0.0, dealiased=dealiased)
    f = (band - none).cpu().numpy()
    expect = np.zeros_like(u)
    for (ix, iy, iz), (_, _, forced) in modes.items():
        if forced:
            expect[:, ix, iy, iz] = coef * u[:, ix, iy, iz]
    assert np.max(np.abs(f - expect)) < 1e-12 * np.max(np.abs(expect))
    assert h3.BandForcing(eps=eps, kmin=2.0, kmax=3.0).injection(d, g) == eps


# --- layout choice without host reads -----------------------------------------------------------


def test_rollout_chain_skips_the_alias_check_and_in_place_edits_void_it(monkeypatch):
    """States the library produced inside the 2/3 box carry a mark, so a chain
    of rollouts never runs the alias check (a device reduction plus a host
    read). Writing into the tensor voids the mark: the next call checks,
    finds content outside the box and takes the dense layout, which keeps the
    reference semantics (those modes feed the nonlinear term)."""
    from hit3d import _lib
    og, g = _grids((32, 32, 32), np.float64)
    calls = []
    real = _lib.call

    def spy(name, *args):
        calls.append(name)
        return real(name, *args)
    monkeypatch.setattr(_lib, "call", spy)     # every module calls through _lib.call
    rhs = h3.make_ns_rhs(g)
    p = h3.NSParams(nu=0.02)
    u0 = h3.random_field(3, g, h3.config_spectrum)
    u1 = h3.rollout(rhs, u0, p, 1e-3, 2)
    u2 = h3.rollout(rhs, u1, p, 1e-3, 2)
    rhs(u2, p, 0.0)
    assert "hit3d_alias_energy" not in calls
    u2[0, 16, 0, 0] = 0.25 + 0.0j           # Nyquist x mode: outside the 2/3 box
    calls.clear()
    u3 = h3.rollout(rhs, u2, p, 1e-3, 1)
    assert calls.count("hit3d_alias_energy") == 1
    ref = O.rollout(O.make_ns_rhs(og), u2.cpu().numpy(), O.NSParams(nu=0.02), 1e-3, 1)
    assert rel(u3, ref) < 1e-12


@pytest.mark.parametrize("shape,dtype", [((64, 64, 64), torch.float64), ((256, 32, 512), torch.float32),
                                         ((16, 128, 8), torch.float64)])
def test_spectrum_box_read_equals_full_read(shape, dtype):
    """A marked (dealiased) state's spectrum reads only the 2/3 box; it
    agrees with the full read of the same data and sums to the energy."""
    g = h3.make_grid(*shape, dtype=dtype)
    u = h3.random_field(4, g, h3.config_spectrum)        # marked by random_field
    box = h3.diagnostics.shell_spectrum(u, g).cpu().numpy()
    plain = h3.diagnostics.shell_spectrum(u.clone(), g).cpu().numpy()   # a clone carries no mark
    for c in range(3):
        assert np.allclose(h3.diagnostics.shell_spectrum(u, g, c).cpu().numpy(),
                           h3.diagnostics.shell_spectrum(u.clone(), g, c).cpu().numpy(), rtol=1e-13, atol=0)
    assert np.allclose(box, plain, rtol=1e-13, atol=0)
    assert abs(box.sum() - h3.energy(u, g)) < 1e-13
    assert np.array_equal(box, h3.diagnostics.shell_spectrum(u, g).cpu().numpy())   # reproducible


# --- integrating-factor (Lawson) RK4 and linear_operator (navierstokes.py:100-103) -----------


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_linear_operator_is_minus_nu_k2(dtype):
    og, g = _grids((32, 16, 8), dtype)
    L = h3.linear_operator(g, h3.NSParams(nu=0.013)).cpu().numpy()
    assert np.array_equal(L, (-0.013 * og.k2).astype(dtype))
    assert np.array_equal(L, O.linear_operator(og, O.NSParams(nu=0.013)).astype(dtype))


@pytest.mark.parametrize("fi", range(len(FORCINGS)))
@pytest.mark.parametrize("dealiased", [True, False])
def test_if_rk4_rollout_matches_oracle(fi, dealiased):
    og, g = _grids((32, 32, 32), np.float64)
    u0 = O.project(random_spectral(50 + fi, og, dealiased), og)
    of, gf = FORCINGS[fi]
    ref = O.rollout(O.make_ns_rhs(og), u0, O.NSParams(nu=0.02, forcing=of()), 4e-3, 6,
                    step=O.if_rk4_step)
    got = h3.rollout(h3.make_ns_rhs(g), dev(u0, g), h3.NSParams(nu=0.02, forcing=gf()), 4e-3, 6,
                     scheme="if-rk4")
    assert rel(got, ref) < 1e-11
    one = h3.if_rk4_step(h3.make_ns_rhs(g), dev(u0, g), h3.NSParams(nu=0.02, forcing=gf()), 4e-3, 0.0)
    assert rel(one, O.if_rk4_step(O.make_ns_rhs(og), u0, O.NSParams(nu=0.02, forcing=of()), 4e-3, 0.0)) < 1e-12


@pytest.mark.parametrize("shape", [(8, 8, 256), (1024, 8, 4)])
def test_if_rk4_large_axis_pruned(shape):
    og, g = _grids(shape, np.float64)
    uh = O.project(random_spectral(14, og), og)
    op = O.NSParams(nu=0.01, forcing=O.LundgrenForcing(nu=0.01, energy=0.5))
    hp = h3.NSParams(nu=0.01, forcing=h3.LundgrenForcing(nu=0.01, energy=0.5))
    ref = O.rollout(O.make_ns_rhs(og), uh, op, 1e-3, 3, step=O.if_rk4_step)
    assert rel(h3.rollout(h3.make_ns_rhs(g), dev(uh, g), hp, 1e-3, 3, scheme="if-rk4"), ref) < 1e-12


def test_if_rk4_beltrami_exact_and_decaying_64_deviation():
    """Beltrami: exact exp(-nu t) decay at a step where RK4 is off. The
    decaying 64^3 case (nu = 1/400, dt = 2e-3, 100 steps, float32): the
    integrating-factor result differs from classic RK4 by its own truncation
    difference, reported here (the reference integrates with RK4)."""
    og, g = _grids((32, 32, 32), np.float64)
    u0 = O.beltrami(og)
    u = h3.rollout(h3.make_ns_rhs(g), dev(u0, g), h3.NSParams(nu=0.5), 0.4, 5, scheme="if-rk4")
    assert rel(u, np.exp(-0.5 * 2.0) * u0) < 1e-13
    g32 = h3.make_grid(64, dtype=torch.float32)
    u0 = h3.random_field(0, g32, h3.config_spectrum)
    rhs, p = h3.make_ns_rhs(g32), h3.NSParams(nu=1 / 400)
    a = h3.rollout(rhs, u0, p, 2e-3, 100)
    b = h3.rollout(rhs, u0, p, 2e-3, 100, scheme="if-rk4")
    dev_ = rel(b.to(torch.complex128), a.to(torch.complex128).cpu().numpy())
    print(f"decaying 64^3, 100 steps: integrating-factor RK4 vs classic RK4 rel L2 = {dev_:.2e}")
    assert dev_ < 1e-4
