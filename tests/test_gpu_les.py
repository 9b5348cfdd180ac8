"""The LES pieces on the GPU: exact adjoints of the differentiable spectral
operators (dot-product identities in float64), the RHS adjoint and a whole
LES step against central differences, and the filtered-DNS subgrid stress
against a NumPy evaluation on the fine grid."""

import numpy as np
import pytest
import torch

import hit3d_oracle as O

pytestmark = pytest.mark.gpu

h3 = pytest.importorskip("hit3d")
from hit3d import _lib, les
from hit3d.closure import ClosureNet
from hit3d.grid import stream


def rdot(a, b):
    """Re sum conj(a) b (torch's pairing for complex tensors)."""
    if a.is_complex():
        return float(torch.sum((a.conj() * b).real))
    return float(torch.sum(a * b))


def crand(shape, gen, dtype=torch.complex128):
    return torch.complex(torch.randn(shape, generator=gen, dtype=torch.float64),
                         torch.randn(shape, generator=gen, dtype=torch.float64)).to(dtype).cuda()


OPS = ["to_physical", "to_spectral", "grad9", "stress_div", "dealias_project"]


@pytest.mark.parametrize("op", OPS)
def test_linear_operator_adjoints(op):
    g = h3.make_grid(16, 8, 16, dtype=torch.float64)
    gen = torch.Generator().manual_seed(OPS.index(op))
    sh = (g.Nx, g.Ny, g.Nz // 2 + 1)
    ph = (g.Nx, g.Ny, g.Nz)
    if op == "to_physical":
        f, x = (lambda x: les._ToPhysical.apply(x, g)), crand((4,) + sh, gen)
    elif op == "to_spectral":
        f, x = (lambda x: les._ToSpectral.apply(x, g)), torch.randn((4,) + ph, generator=gen,
                                                                     dtype=torch.float64).cuda()
    elif op == "grad9":
        f, x = (lambda x: les._Grad9.apply(x, g)), crand((3,) + sh, gen)
    elif op == "stress_div":
        f, x = (lambda x: les._StressDiv.apply(x, g)), crand((6,) + sh, gen)
    else:
        f, x = (lambda x: les._DealiasProject.apply(x, g)), crand((3,) + sh, gen)
    x = x.requires_grad_(True)
    y = f(x)
    w = crand(y.shape, gen) if y.is_complex() else torch.randn(y.shape, generator=gen,
                                                               dtype=torch.float64).cuda()
    (gx,) = torch.autograd.grad(y, x, w)
    lhs, rhs = rdot(w, y.detach()), rdot(gx, x.detach())
    scale = float(torch.sum(w.abs() * y.detach().abs()))
    assert abs(lhs - rhs) <= 1e-13 * scale


def test_ns_rhs_backward_matches_central_differences():
    g = h3.make_grid(16, dtype=torch.float64)
    rhs = h3.make_ns_rhs(g)
    params = h3.NSParams(nu=0.02, forcing=h3.LinearForcing(A=0.1))
    u = h3.random_field(1, g, h3.config_spectrum).requires_grad_(True)
    gen = torch.Generator().manual_seed(2)
    W = crand(u.shape, gen)
    y = les._NSRhs.apply(u, rhs, params, 0.0)
    (gu,) = torch.autograd.grad(y, u, W)
    d = h3.project(h3.dealias(crand(u.shape, gen), g), g)
    eps = 1e-4 * float(u.detach().abs().max() / d.abs().max())

    def J(v):
        return rdot(W, rhs(v, params, 0.0))
    fd = (J(u.detach() + eps * d) - J(u.detach() - eps * d)) / (2 * eps)
    assert abs(rdot(gu, d) - fd) < 1e-7 * abs(fd)


def test_les_step_gradient_without_closure_matches_central_differences():
    g = h3.make_grid(16, dtype=torch.float64)
    rhs = h3.make_ns_rhs(g)
    params = h3.NSParams(nu=0.02, forcing=h3.LinearForcing(A=0.1))
    u0 = h3.random_field(3, g, h3.config_spectrum)
    gen = torch.Generator().manual_seed(4)
    W = crand(u0.shape, gen)

    def J(v):
        u = v
        for n in range(3):
            u = les.les_step(u, rhs, params, None, 5e-3, n * 5e-3)
        return torch.sum((W.conj() * u).real)
    x = u0.clone().requires_grad_(True)
    (gx,) = torch.autograd.grad(J(x), x)
    d = h3.project(h3.dealias(crand(u0.shape, gen), g), g)
    eps = 1e-4 * float(u0.abs().max() / d.abs().max())
    fd = (float(J(u0 + eps * d)) - float(J(u0 - eps * d))) / (2 * eps)
    assert abs(rdot(gx, d) - fd) < 1e-7 * abs(fd)


def test_les_step_with_closure_gradient_matches_central_differences():
    """float32 LES with the bf16 net: the autograd gradient of a functional of
    two steps against central differences (float32 + bf16 noise: 2%)."""
    g = h3.make_grid(16, dtype=torch.float32)
    rhs = h3.make_ns_rhs(g)
    params = h3.NSParams(nu=0.02, forcing=h3.LinearForcing(A=0.1))
    net = ClosureNet(g, channels=16, seed=1)
    net.clip_backscatter = False      # the clipping mask is not differentiable across its edge
    u0 = h3.random_field(5, g, h3.config_spectrum)
    gg = les.velocity_gradient(u0, g).reshape(9, -1)
    net.set_scales(1.0 / gg.std(dim=1), torch.full((6,), 0.05))
    gen = torch.Generator().manual_seed(6)
    W = crand(u0.shape, gen, torch.complex64)

    def J(v):
        u = v
        for n in range(2):
            u = les.les_step(u, rhs, params, net, 5e-3, n * 5e-3)
        return torch.sum((W.conj() * u).real.double())
    x = u0.clone().requires_grad_(True)
    net.zero_grad()
    (gx,) = torch.autograd.grad(J(x), x)
    d = h3.project(h3.dealias(crand(u0.shape, gen, torch.complex64), g), g)
    eps = 3e-3 * float(u0.abs().max() / d.abs().max())
    with torch.no_grad():
        fd = (float(J(u0 + eps * d)) - float(J(u0 - eps * d))) / (2 * eps)
    assert abs(rdot(gx.to(torch.complex128), d.to(torch.complex128)) - fd) < 2e-2 * abs(fd)
    assert float(net.layers[2].gw.abs().max()) > 0        # the closure received gradient


def test_subgrid_stress_matches_numpy():
    fine, coarse = h3.make_grid(32, dtype=torch.float64), h3.make_grid(16, dtype=torch.float64)
    ofine = O.make_grid(32, dtype=np.float64)
    u = O.random_field(7, ofine, O.config_spectrum)
    ubar, tau_hat = les.subgrid_stress(torch.from_numpy(u).cuda(), fine, coarse, 5.0)
    # NumPy on the fine grid: G = sharp filter |k| <= 5, then sample every 2nd point
    mask = ofine.k2 <= 25.0
    G = lambda f: O.to_physical(O.to_spectral(f[None], ofine) * mask, ofine)[0]
    up = O.to_physical(u, ofine)
    ub = O.to_physical(u * mask, ofine)
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    tau = np.stack([G(up[i] * up[j]) - G(ub[i] * ub[j]) for i, j in pairs])[:, ::2, ::2, ::2]
    got = h3.to_physical(tau_hat, coarse).cpu().numpy()
    assert np.abs(got - tau).max() < 1e-12 * np.abs(tau).max()
    assert np.abs(h3.to_physical(ubar, coarse).cpu().numpy() - ub[:, ::2, ::2, ::2]).max() < 1e-12


def test_les_experiment_driver_end_to_end(tmp_path):
    """The whole pipeline at toy size: DNS 64^3 -> LES 16^3 (kcut 5), data,
    a-priori and a-posteriori training, evaluation and timing."""
    from hit3d import les_experiment
    res = les_experiment.run(fine_n=64, les_n=16, kcut=5.0, nu=1 / 200, dt_dns=4e-3, stride=2,
                             spinup_time=0.2, nsnap=24, holdout_time=0.4, epochs=1, windows=2, window=3,
                             attractor_windows=2, spin=3, calib_steps=20, channels=16,
                             save=str(tmp_path / "closure.npz"))
    assert res["config"]["eval_steps"] == 49 and len(res["les"]["E_les"]) == 5
    assert res["config"]["holdout_snapshots"] == 50 and res["acceptance_statistic"] == "held-out"
    assert res["les"]["long_steps"] == 147 and np.isfinite(res["les"]["max_rel_err_long"])
    assert res["les_in_sample"]["steps"] == 23 and np.isfinite(res["les_in_sample"]["max_rel_err"])
    assert np.isfinite(res["training_half_vs_half_max"]) and np.isfinite(res["training_vs_holdout_max"])
    net, target, e_target, config = les_experiment.load_closure(str(tmp_path / "closure.npz"),
                                                                h3.make_grid(16, dtype=torch.float32))
    assert len(target) == 5 and e_target > 0 and config["kcut"] == 5.0
    assert abs(float(net.gain.detach()) - res["closure_gain"]) < 1e-6 and net.tilt == res["closure_tilt"]
    assert res["config"]["forcing_les"].startswith("HoldResolvedEnergy")
    assert len(res["closure_calibration"]) == 7          # start + 2 x (2 differences + step)
    best = min(res["closure_calibration"], key=lambda c: c["rms_log_error"])
    assert abs(res["closure_gain"] - best["gain"]) <= 1e-6 and res["closure_tilt"] == best["tilt"]
    assert all(np.isfinite(res["les"]["rel_err"])) and all(np.isfinite(res["les_no_closure"]["rel_err"]))
    assert res["gpu_seconds"]["dns"] > 0 and res["gpu_seconds"]["les"] > 0
    assert isinstance(res["accepted"], bool)
    assert res["fallback"] == ("none" if res["accepted"] else "plain DNS (closure error above tolerance)")


def test_backscatter_clipping_keeps_transfer_dissipative():
    gen = torch.Generator().manual_seed(9)
    g = torch.randn(9, 4096, generator=gen).cuda()
    tau = torch.randn(6, 4096, generator=gen).cuda()
    out = les.clip_backscatter(tau, g)
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]

    def transfer(t):
        return -sum((1.0 if i == j else 2.0) * t[s] * 0.5 * (g[3 * j + i] + g[3 * i + j])
                    for s, (i, j) in enumerate(pairs))
    pi_in, pi_out = transfer(tau), transfer(out)
    assert float(pi_out.min()) >= 0.0
    keep = pi_in >= 0
    assert torch.equal(out[:, keep], tau[:, keep]) and float(out[:, ~keep].abs().max()) == 0.0
    # the library's fused unpack + clip against this statement (raw net output
    # aos[v][c] scaled by s[c]); the kept/zeroed decision may differ only where
    # Pi is at rounding level
    s = torch.linspace(0.5, 1.5, 6).cuda()
    aos = (tau / s[:, None]).t().contiguous()
    got = torch.empty(6, 4096, device="cuda")
    kept = torch.empty(4096, dtype=torch.uint8, device="cuda")
    h3._lib.call("hit3d_stress_unpack_clip", aos.data_ptr(), 4096, s.data_ptr(), g.data_ptr(),
                 got.data_ptr(), kept.data_ptr(), h3.grid.stream())
    ref = les.clip_backscatter((aos * s[None, :]).t().contiguous(), g)
    diff = (kept.bool() != (transfer((aos * s[None, :]).t()) >= 0))
    scale = float(transfer((aos * s[None, :]).t()).abs().max())
    assert float(transfer((aos * s[None, :]).t())[diff].abs().max() if diff.any() else 0.0) <= 1e-5 * scale
    same = ~diff
    assert float((got[:, same] - ref[:, same]).abs().max()) <= 1e-6 * float(ref.abs().max())


def test_les_graph_replays_match_eager_steps():
    """A captured LES step (closure, spherical mask) replayed several times,
    with eager work and synchronisation between replays; every replay equals
    the eager step from the same state."""
    g = h3.make_grid(32, dtype=torch.float32)
    rhs = h3.make_ns_rhs(g)
    params = h3.NSParams(nu=0.01, forcing=h3.LinearForcing(A=0.1))
    net = ClosureNet(g, channels=32, seed=2)
    u0 = les.filter_state(h3.random_field(4, g, h3.config_spectrum), g, g, 10.0)
    gg = les.velocity_gradient(u0, g).reshape(9, -1)
    net.set_scales(1.0 / gg.std(dim=1), torch.full((6,), 0.05))
    mask = les.sphere_mask(g, 10.0)
    G = les.LesStepGraph(u0, rhs, params, net, 4e-3, mask)
    u_eager = u0.clone()
    for r in range(4):
        with torch.no_grad():
            u_eager = les.les_step(u_eager, rhs, params, net, 4e-3, 0.0, mask)
        h3.energy(u_eager, g)                     # eager work and a host sync between replays
        torch.cuda.synchronize()
        out = G.step().clone()
        assert torch.equal(out, u_eager), r
    assert float((u_eager * (1 - mask)).abs().max()) == 0.0   # stays inside the sphere


def _hold_reference(u, du, g, energy, tau):
    """float64 torch statement of hit3d_hold_energy."""
    w = g.weights.double() / float(g.Nx * g.Ny * g.Nz) ** 2
    u, du = u.to(torch.complex128), du.to(torch.complex128)
    s1 = torch.sum(w * (u.real * du.real + u.imag * du.imag))
    s2 = torch.sum(w * (u.real ** 2 + u.imag ** 2))
    rate = -s1 - ((0.5 * s2 - energy) / tau if tau else 0.0)
    A = rate / s2
    return du + A * u, torch.stack([A, s1, s2])


@pytest.mark.parametrize("dtype,tau", [(torch.float64, 0.5), (torch.float64, None), (torch.float32, 1.0)])
def test_hold_energy_matches_formula_and_sets_the_energy_rate(dtype, tau):
    g = h3.make_grid(16, 8, 16, dtype=dtype)
    gen = torch.Generator().manual_seed(11)
    cd = torch.complex128 if dtype == torch.float64 else torch.complex64
    u = h3.dealias(crand((3, g.Nx, g.Ny, g.Nz // 2 + 1), gen, cd), g)
    du = crand(u.shape, gen, cd)
    out, coef = les._hold_raw(u, du, g, 0.3, tau)
    ref, rc = _hold_reference(u, du, g, 0.3, tau)
    tol = 1e-12 if dtype == torch.float64 else 2e-6
    assert torch.allclose(coef, rc, rtol=tol, atol=0), (coef, rc)
    assert float((out.to(torch.complex128) - ref).abs().max()) <= tol * float(ref.abs().max())
    # the held rate: dE/dt of the output = -(E - energy)/tau (0 without relaxation)
    w = g.weights.double() / float(g.Nx * g.Ny * g.Nz) ** 2
    o, uu = out.to(torch.complex128), u.to(torch.complex128)
    dE = float(torch.sum(w * (uu.real * o.real + uu.imag * o.imag)))
    E = 0.5 * float(rc[2])
    want = -(E - 0.3) / tau if tau else 0.0
    scale = float(torch.sum(w * uu.abs() * o.abs()))
    assert abs(dE - want) <= (1e-13 if dtype == torch.float64 else 1e-6) * scale


def test_hold_energy_backward_matches_autograd_of_the_formula():
    g = h3.make_grid(16, dtype=torch.float64)
    gen = torch.Generator().manual_seed(12)
    u = h3.dealias(crand((3, g.Nx, g.Ny, g.Nz // 2 + 1), gen), g).requires_grad_(True)
    du = crand(u.shape, gen).requires_grad_(True)
    W = crand(u.shape, gen)
    for tau in (0.7, None):
        gu, gdu = torch.autograd.grad(les.hold_energy(u, du, g, 0.2, tau), (u, du), W)
        ru, rdu = torch.autograd.grad(_hold_reference(u, du, g, 0.2, tau)[0], (u, du), W)
        assert float((gu - ru).abs().max()) <= 1e-12 * float(ru.abs().max())
        assert float((gdu - rdu).abs().max()) <= 1e-12 * float(rdu.abs().max())


def test_hold_energy_rejects_bad_arguments():
    g = h3.make_grid(8, dtype=torch.float32)
    u = torch.zeros((3, 8, 8, 5), dtype=torch.complex64, device="cuda")
    with pytest.raises(ValueError):
        les._hold_raw(u, u, g, -1.0, 1.0)


def test_les_graph_with_hold_energy_matches_eager_and_holds_the_energy():
    """The energy-holding forcing inside a captured LES step: replays equal
    eager steps, and without relaxation the resolved energy stays at its
    start value up to the RK4 error of a conserved quadratic."""
    g = h3.make_grid(32, dtype=torch.float32)
    rhs = h3.make_ns_rhs(g)
    net = ClosureNet(g, channels=32, seed=3)
    u0 = les.filter_state(h3.random_field(6, g, h3.config_spectrum), g, g, 10.0)
    gg = les.velocity_gradient(u0, g).reshape(9, -1)
    net.set_scales(1.0 / gg.std(dim=1), torch.full((6,), 0.05))
    mask = les.sphere_mask(g, 10.0)
    E0 = h3.energy(u0, g)
    params = h3.NSParams(nu=0.01, forcing=les.HoldResolvedEnergy(E0, None))
    G = les.LesStepGraph(u0, rhs, params, net, 4e-3, mask)
    u_eager = u0.clone()
    for r in range(4):
        with torch.no_grad():
            u_eager = les.les_step(u_eager, rhs, params, net, 4e-3, 0.0, mask)
        h3.energy(u_eager, g)
        torch.cuda.synchronize()
        out = G.step().clone()
        assert torch.equal(out, u_eager), r
    assert abs(h3.energy(u_eager, g) / E0 - 1) < 1e-5


def test_fused_les_rhs_matches_the_differentiable_composition():
    """Without autograd, les_rhs assembles the closure force, the sum and the
    spherical mask in one library pass (hit3d_les_combine); it must equal the
    differentiable composition (stress divergence, P M, gain and spectral
    shape, mask) that training uses, with and without the closure."""
    g = h3.make_grid(32, dtype=torch.float32)
    fine = h3.make_grid(64, dtype=torch.float32)
    u = les.filter_state(h3.random_field(3, fine, h3.config_spectrum), fine, g, 10.0)
    rhs = h3.make_ns_rhs(g)
    net = ClosureNet(g, channels=16, seed=1)
    net.set_tilt(-0.2)
    with torch.no_grad():
        net.gain.fill_(1.3)
    mask = les.sphere_mask(g, 10.0)
    params = h3.NSParams(nu=1 / 500, forcing=les.HoldResolvedEnergy(0.3, 1.0))
    for model in (net, None):
        with torch.enable_grad():
            ref = les.les_rhs(u, rhs, params, model, 0.0, mask).detach()
        with torch.no_grad():
            got = les.les_rhs(u, rhs, params, model, 0.0, mask)
        err = float(torch.linalg.vector_norm(got - ref) / torch.linalg.vector_norm(ref))
        assert err < 2e-6, (model is None, err)
        assert float(torch.abs(got * (1 - mask)).max()) == 0.0


def test_closure_spectral_shape_uses_the_physical_cut():
    """S(k) = exp(tilt (|k|/kc - 1/2)) with kc = (2 pi / L)(N // 3): on a box
    of side pi (k0 = 2) at N = 32 the mode n = (5, 0, 0) has |k| = 10 = kc/2,
    where S = 1 for any tilt, and n = (10, 0, 0) has S = exp(tilt / 2)."""
    g = h3.make_grid(32, Lx=np.pi, Ly=np.pi, Lz=np.pi, dtype=torch.float32)
    net = ClosureNet(g, channels=16, seed=0)
    net.set_tilt(0.4)
    S = net.spectral_shape().cpu().numpy()
    assert abs(S[5, 0, 0] - 1.0) < 1e-6
    assert abs(S[10, 0, 0] - np.exp(0.2)) < 1e-6


GOLDEN_CLOSURE = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                                            "les_closure_64.npz")


def _les_from_target(net_or_none, target, e_target, steps, dt, seed=5):
    """64^3 LES (|k| <= 21, nu = 1/2000, step dt, resolved energy held at the
    training record's mean) from a random-phase field carrying the target
    spectrum; returns the spectra (shells 1..21) every 10 steps."""
    from hit3d.les_experiment import _spectrum
    g = h3.make_grid(64, dtype=torch.float32)
    spec = lambda k: np.where((k >= 0.5) & (k <= 21.5), np.interp(k, np.arange(1, 22), target), 0.0)  # noqa: E731
    u = les.filter_state(h3.random_field(seed, g, spec), g, g, 21.0)
    params = h3.NSParams(nu=1 / 2000, forcing=les.HoldResolvedEnergy(e_target, 1.0))
    out = []
    les.les_rollout(u, h3.make_ns_rhs(g), params, net_or_none, dt, steps, every=10,
                    mask=les.sphere_mask(g, 21.0), observe=lambda n, v: out.append(_spectrum(v, g, 21)))
    return np.asarray(out)


def test_les_with_the_committed_closure_reproduces_the_filtered_dns_spectrum():
    """Regression of the trained closure (tests/golden/les_closure_64.npz: the
    weights, gain and spectral tilt of the 512^3 -> 64^3 experiment, with the
    filtered DNS's mean spectrum of its training record): an LES from a
    random-phase field holds the filtered DNS spectrum; without the closure
    the spectrum piles up at the cut."""
    from hit3d.les_experiment import load_closure
    g = h3.make_grid(64, dtype=torch.float32)
    net, target, e_target, config = load_closure(GOLDEN_CLOSURE, g)
    assert config["les_n"] == 64 and config["kcut"] == 21.0 and len(target) == 21
    dt = config["dt_les"]
    n = int(round(20.0 / dt))                     # 20 time units, the first 5 discarded
    spectra = _les_from_target(net, target, e_target, n, dt)
    E = spectra[n // 40:].mean(0)
    err = np.abs(E / target - 1.0)
    bare = _les_from_target(None, target, e_target, n // 4, dt)[n // 80:].mean(0)
    err_bare = np.abs(bare / target - 1.0)
    print(f"LES with the committed closure vs the filtered DNS mean: max rel err k=1..2 "
          f"{err[:2].max():.3f}, k=3..21 {err[2:].max():.3f} (k={int(np.argmax(err[2:])) + 3}); "
          f"without closure at k=21: {err_bare[-1]:.2f}")
    assert np.all(np.isfinite(E))
    # measured on an MI355X: 0.113 (k = 3..21) and 0.120 (k = 1, 2)
    assert err[2:].max() < 0.17 and err[:2].max() < 0.18
    assert err_bare[-1] > 1.0 and err_bare[-1] > 5 * err[-1]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_rk4_combine_matches_the_stage_sums(dtype):
    """hit3d_rk4_combine against classic RK4's stage sums (integrators.py:32-38)
    evaluated in float64 from the same inputs; a wrong stage coefficient, a
    stage 2/3 that restarts acc, or a stage 4 that reads u fails."""
    g = h3.make_grid(16, dtype=dtype)
    gen = torch.Generator().manual_seed(5)
    shape = (3, 16, 16, 9)
    cd = g.complex_dtype
    u, k, acc0 = (crand(shape, gen, cd) for _ in range(3))
    dt = 0.37
    tol = 1e-6 if dtype == torch.float32 else 1e-15
    want = {1: (u + dt / 2 * k, u + dt / 6 * k), 2: (u + dt / 2 * k, acc0 + dt / 3 * k),
            3: (u + dt * k, acc0 + dt / 3 * k), 4: (acc0 + dt / 6 * k, acc0)}
    for stage, (w_out, w_acc) in want.items():
        acc = acc0.clone()
        out = torch.empty_like(u)
        _lib.call("hit3d_rk4_combine", g.desc, stage, dt, None if stage == 4 else u.data_ptr(),
                  k.data_ptr(), None, None, acc.data_ptr(), out.data_ptr(), stream())
        for got, ref in ((out, w_out), (acc, w_acc)):
            err = float((got.to(torch.complex128) - ref.to(torch.complex128)).abs().max())
            assert err <= tol * float(ref.abs().max()), (stage, err)
    # with the energy hold's term: the derivative is k + coef[0] us
    us = crand(shape, gen, cd)
    coef = torch.tensor([0.7, 0.0, 0.0], dtype=torch.float64, device="cuda")
    acc = acc0.clone()
    out = torch.empty_like(u)
    _lib.call("hit3d_rk4_combine", g.desc, 2, dt, u.data_ptr(), k.data_ptr(), coef.data_ptr(),
              us.data_ptr(), acc.data_ptr(), out.data_ptr(), stream())
    kh = k.to(torch.complex128) + 0.7 * us.to(torch.complex128)
    for got, ref in ((out, u.to(torch.complex128) + dt / 2 * kh), (acc, acc0.to(torch.complex128) + dt / 3 * kh)):
        err = float((got.to(torch.complex128) - ref).abs().max())
        assert err <= 2 * tol * float(ref.abs().max()), err
    with pytest.raises(ValueError):
        _lib.call("hit3d_rk4_combine", g.desc, 5, dt, u.data_ptr(), k.data_ptr(), None, None,
                  acc.data_ptr(), out.data_ptr(), stream())
    with pytest.raises(ValueError):
        _lib.call("hit3d_rk4_combine", g.desc, 2, dt, u.data_ptr(), k.data_ptr(), None, None,
                  acc.data_ptr(), acc.data_ptr(), stream())


def test_les_step_without_gradients_matches_the_tensor_op_step():
    """The no-gradient LES step (fused right-hand side, one combination pass
    per stage, new state written over u) against the differentiable step built
    from tensor ops: three steps without the closure, and one step with it.
    With the closure the two forms differ by more than round-off: their stage
    inputs differ in the last bit (one rounding in the combination pass, two
    in the tensor ops), and the BF16 rounding of the closure's input turns
    that into a difference of one BF16 unit (2^-8) at some voxels of an
    untrained net's force (7.9e-5 relative after one step on an MI355X); a
    wrong stage coefficient or a dropped stage moves the state by O(dt)."""
    g = h3.make_grid(32, dtype=torch.float32)
    fine = h3.make_grid(64, dtype=torch.float32)
    u0 = les.filter_state(h3.random_field(4, fine, h3.config_spectrum), fine, g, 10.0)
    rhs = h3.make_ns_rhs(g)
    net = ClosureNet(g, channels=16, seed=2)
    mask = les.sphere_mask(g, 10.0)
    params = h3.NSParams(nu=1 / 500, forcing=les.HoldResolvedEnergy(0.3, 1.0))

    def rel(a, b):
        return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))

    ref, got = u0.clone(), u0.clone()
    for n in range(3):
        with torch.enable_grad():
            ref = les.les_step(ref, rhs, params, None, 4e-3, n * 4e-3, mask).detach()
        with torch.no_grad():
            same = les.les_step(got, rhs, params, None, 4e-3, n * 4e-3, mask, out=got)
        assert same.data_ptr() == got.data_ptr()
    assert rel(got, ref) < 5e-6, rel(got, ref)
    assert rel(got, u0) > 1e-3                     # the steps moved the state
    with torch.enable_grad():
        ref = les.les_step(u0, rhs, params, net, 4e-3, 0.0, mask).detach()
    with torch.no_grad():
        got = les.les_step(u0, rhs, params, net, 4e-3, 0.0, mask)
    assert rel(got, ref) < 3e-4, rel(got, ref)
    with torch.enable_grad(), pytest.raises(ValueError):
        les.les_step(got, rhs, params, net, 4e-3, 0.0, mask, out=got)
