"""The hand-written adjoint on the GPU: parity with the oracle's adjoint
(itself pinned by central differences in test_oracle.py), the autograd
wrapper, and a central-difference check of the adjoint configuration's
objective on the GPU."""

import numpy as np
import pytest
import torch

import hit3d_oracle as O

pytestmark = pytest.mark.gpu

h3 = pytest.importorskip("hit3d")

FORCINGS = [
    (O.NoForcing(), h3.NoForcing()),
    (O.LinearForcing(A=0.15), h3.LinearForcing(A=0.15)),
    (O.BandForcing(eps=0.3, kmin=1.0, kmax=3.0), h3.BandForcing(eps=0.3, kmin=1.0, kmax=3.0)),
    (O.LundgrenForcing(nu=0.02), h3.LundgrenForcing(nu=0.02)),
    (O.LundgrenForcing(nu=0.02, energy=0.6, tau=0.3), h3.LundgrenForcing(nu=0.02, energy=0.6, tau=0.3)),
]


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(np.asarray(b).ravel()))


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _rand_c(seed, shape):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("fi", range(len(FORCINGS)))
@pytest.mark.parametrize("shape", [(16, 16, 16), (32, 16, 8)])
@pytest.mark.parametrize("dealiased", [False, True])
def test_rhs_vjp_matches_oracle(fi, shape, dealiased):
    """Dense input (content outside the 2/3 box) through the dense
    transforms; a dealiased state through the pruned inverse transforms."""
    og, g = O.make_grid(*shape), h3.make_grid(*shape, dtype=torch.float64)
    u = O.random_field(1, og, O.config_spectrum)
    if not dealiased:
        u = u + 1e-3 * np.abs(u).max() * _rand_c(2, u.shape)
    v = _rand_c(3, u.shape)
    of, gf = FORCINGS[fi]
    ref = O.make_ns_rhs_vjp(og)(u, v, O.NSParams(nu=0.02, forcing=of), 0.0)
    got = h3.make_ns_rhs_vjp(g)(dev(u), dev(v), h3.NSParams(nu=0.02, forcing=gf), 0.0, dealiased=dealiased)
    assert rel(got, ref) < 1e-12


@pytest.mark.parametrize("fi", [0, 2, 3])
def test_rk4_step_and_rollout_vjp_match_oracle(fi):
    og, g = O.make_grid(16), h3.make_grid(16, dtype=torch.float64)
    u = O.random_field(4, og, O.config_spectrum)
    ub = _rand_c(5, u.shape)
    of, gf = FORCINGS[fi]
    op, gp = O.NSParams(nu=0.02, forcing=of), h3.NSParams(nu=0.02, forcing=gf)
    orhs, grhs = O.make_ns_rhs(og), h3.make_ns_rhs(g)
    ref = O.rk4_step_vjp(orhs, O.make_ns_rhs_vjp(og), u, ub, op, 2e-3, 0.0)
    assert rel(h3.rk4_step_vjp(grhs, dev(u), dev(ub), gp, 2e-3, 0.0), ref) < 1e-12
    ref = O.rollout_vjp(orhs, O.make_ns_rhs_vjp(og), u, ub, op, 2e-3, 7, checkpoint=3)
    for ck in (1, 3, 20):
        assert rel(h3.rollout_vjp(grhs, dev(u), dev(ub), gp, 2e-3, 7, checkpoint=ck), ref) < 1e-11


@pytest.mark.parametrize("dealiased", [True, False])
def test_rk4_steps_record_keeps_states_and_stage_inputs(dealiased):
    """hit3d_rk4_steps_record: slot 0 of step n is the state after n steps
    (bit for bit the rollout's), slots 1..3 the stage inputs of the oracle's
    RK4 step (integrators.py:35-37) from that state; dense inputs included."""
    import ctypes
    from hit3d import _lib
    from hit3d.grid import stream
    og, g = O.make_grid(16), h3.make_grid(16, dtype=torch.float64)
    u = O.random_field(8, og, O.config_spectrum)
    if not dealiased:
        u = u + 1e-3 * np.abs(u).max() * _rand_c(9, u.shape)
    of, gf = FORCINGS[4]
    op, gp = O.NSParams(nu=0.02, forcing=of), h3.NSParams(nu=0.02, forcing=gf)
    dt, n = 2e-3, 3
    rec = torch.empty((n, 4) + u.shape, dtype=torch.complex128, device="cuda")
    x = dev(u)
    pruned = 1 if dealiased else 0
    ws = g.workspace(_lib.OP_RK4, 3, pruned)
    f = gf.desc()
    _lib.call("hit3d_rk4_steps_record", g.desc, x.data_ptr(), 0.02, dt, n, 0.0, ctypes.byref(f), pruned,
              rec.data_ptr(), ws.data_ptr(), ws.numel(), stream())
    assert torch.equal(x, h3.rollout(h3.make_ns_rhs(g), dev(u), gp, dt, n))
    orhs, un = O.make_ns_rhs(og), u
    for step in range(n):
        got = rec[step].cpu().numpy()
        assert torch.equal(rec[step, 0], h3.rollout(h3.make_ns_rhs(g), dev(u), gp, dt, step))
        k1 = orhs(un, op, 0.0)
        y2 = un + 0.5 * dt * k1
        k2 = orhs(y2, op, 0.0)
        y3 = un + 0.5 * dt * k2
        y4 = un + dt * orhs(y3, op, 0.0)
        for slot, ref in ((1, y2), (2, y3), (3, y4)):
            assert rel(got[slot], ref) < 1e-13
        un = O.rollout(orhs, un, op, dt, 1)


@pytest.mark.parametrize("shape", [(8, 8, 8), (32, 16, 8), (16, 8, 32), (8, 4, 256), (256, 8, 4)])
def test_step_vjp_pruned_transforms_equal_dense(shape):
    """The step adjoint on a dealiased state takes the pruned inverse
    transforms (compact prep, pruned x/y passes, z pass on kz-mode pencils)
    when the state is marked; an unmarked copy takes the dense ones. Small
    grids (row padding makes the pruned layouts larger than a dense field)
    and long axes included."""
    from hit3d.grid import mark_dealiased
    g = h3.make_grid(*shape, dtype=torch.float64)
    u = h3.random_field(12, g, h3.config_spectrum)
    ub = dev(_rand_c(13, tuple(u.shape)))
    params = h3.NSParams(nu=0.02, forcing=h3.LundgrenForcing(nu=0.02))
    rhs = h3.make_ns_rhs(g)
    marked = mark_dealiased(u.clone())
    plain = u.clone()
    a = h3.rk4_step_vjp(rhs, marked, ub, params, 2e-3, 0.0)
    b = h3.rk4_step_vjp(rhs, plain, ub, params, 2e-3, 0.0)
    assert rel(a, b.cpu().numpy()) < 1e-13


def test_rollout_vjp_record_and_recompute_agree(monkeypatch):
    """The recorded segments (stage inputs kept by the forward recompute) and
    the recompute-per-step path (taken when a record does not fit) give the
    same cotangent and parameter gradients."""
    import hit3d.adjoint as A
    g = h3.make_grid(32, dtype=torch.float64)
    u = h3.random_field(10, g, h3.config_spectrum)
    ub = dev(_rand_c(11, tuple(u.shape)))
    params = h3.NSParams(nu=0.02, forcing=h3.BandForcing(eps=0.3, kmin=1.0, kmax=3.0))
    rhs = h3.make_ns_rhs(g)
    pg1, pg2 = A.param_grad_buffer(u.device), A.param_grad_buffer(u.device)
    a = h3.rollout_vjp(rhs, u, ub, params, 2e-3, 9, checkpoint=4, param_grad=pg1)
    monkeypatch.setattr(A, "_RECORD_SHARE", 0.0)
    b = h3.rollout_vjp(rhs, u, ub, params, 2e-3, 9, checkpoint=4, param_grad=pg2)
    assert rel(a, b.cpu().numpy()) < 1e-13
    assert float((pg1 - pg2).abs().max() / pg2.abs().max()) < 1e-12


def _shell_energy_weights(g, lo, hi):
    """w / (2 Ntot^2) on modes whose shell s has lo <= s < hi (the energy in
    those shells is sum(W |u|^2))."""
    shell = g.shell.to(torch.int64)
    band = ((shell >= lo) & (shell < hi)).to(g.dtype)
    return band * g.weights / (2.0 * float(g.Nx * g.Ny * g.Nz) ** 2)


def test_autograd_through_rollout_matches_oracle():
    og, g = O.make_grid(16), h3.make_grid(16, dtype=torch.float64)
    u = O.random_field(6, og, O.config_spectrum)
    params = h3.NSParams(nu=0.02, forcing=h3.BandForcing(eps=0.1, kmin=1.0, kmax=2.0))
    W = _shell_energy_weights(g, 3, 6)
    u0 = dev(u).requires_grad_(True)
    uN = h3.differentiable_rollout(h3.make_ns_rhs(g), u0, params, 2e-3, 6, checkpoint=4)
    J = torch.sum(W * (uN.real ** 2 + uN.imag ** 2))
    J.backward()
    # dJ/du_N = 2 W u_N (torch convention); the oracle maps it back through the steps
    op = O.NSParams(nu=0.02, forcing=O.BandForcing(eps=0.1, kmin=1.0, kmax=2.0))
    uNo = O.rollout(O.make_ns_rhs(og), u, op, 2e-3, 6)
    ref = O.rollout_vjp(O.make_ns_rhs(og), O.make_ns_rhs_vjp(og), u, 2 * W.cpu().numpy() * uNo, op,
                        2e-3, 6, checkpoint=4)
    assert rel(u0.grad, ref) < 1e-11


def _adjoint_config_grad(dtype, N, steps, u0=None):
    """The adjoint configuration: J = energy in shells 8 <= |k| < 16 after
    `steps` RK4 steps (dt = 2e-3, nu = 1/600, band forcing eps = 0.1 on
    1 <= |k| <= 2), gradient by the checkpointed adjoint (every 20 steps)."""
    g = h3.make_grid(N, dtype=dtype)
    params = h3.NSParams(nu=1 / 600, forcing=h3.BandForcing(eps=0.1, kmin=1.0, kmax=2.0))
    rhs = h3.make_ns_rhs(g)
    if u0 is None:
        u0 = h3.random_field(0, g, h3.config_spectrum)
    W = _shell_energy_weights(g, 8, 16)
    x = u0.clone().requires_grad_(True)
    uN = h3.differentiable_rollout(rhs, x, params, 2e-3, steps, checkpoint=20)
    torch.sum(W * (uN.real ** 2 + uN.imag ** 2)).backward()
    return g, params, rhs, u0, W, x.grad.detach()


def test_adjoint_config_float32_gradient_matches_float64():
    """128^3, 200 steps: the float32 adjoint gradient against the float64 one
    from the same initial field (relative L2)."""
    g32, _, _, u32, _, g_32 = _adjoint_config_grad(torch.float32, 128, 200)
    _, _, _, _, _, g_64 = _adjoint_config_grad(torch.float64, 128, 200, u32.to(torch.complex128))
    err = float(torch.linalg.vector_norm(g_32.to(torch.complex128) - g_64) / torch.linalg.vector_norm(g_64))
    print(f"adjoint config: float32 vs float64 gradient rel L2 = {err:.2e}")
    assert err < 1e-4


@pytest.mark.parametrize("dtype,N,steps", [(torch.float64, 32, 20), (torch.float64, 128, 200)])
def test_adjoint_objective_central_differences(dtype, N, steps):
    """The adjoint configuration's gradient against central differences of J
    along 16 seeded random solenoidal directions, eps = 1e-3 of the field's
    scale; 128^3 x 200 steps is the configuration itself. A central
    difference carries an O(eps^2) truncation error (bound 5e-5 relative);
    the Richardson combination (4 D(eps/2) - D(eps)) / 3 removes it to
    O(eps^4), which the adjoint must meet to 1e-8. The pairing
    an = Re sum(conj(grad) * d) is torch's convention; jax's cotangent for
    the same J is conj(grad)."""
    g, params, rhs, u0, W, grad = _adjoint_config_grad(dtype, N, steps)

    def J(u):
        uN = h3.rollout(rhs, u, params, 2e-3, steps)
        return float(torch.sum(W * (uN.real.double() ** 2 + uN.imag.double() ** 2)))

    scale = float(torch.sqrt(torch.mean(u0.real ** 2 + u0.imag ** 2)))
    gen = torch.Generator(device="cpu").manual_seed(1)
    err_cd, err_rich = [], []
    for _ in range(16):
        d = torch.complex(torch.randn(u0.shape, generator=gen, dtype=torch.float64),
                          torch.randn(u0.shape, generator=gen, dtype=torch.float64))
        d = h3.project(h3.dealias(d.to(u0.dtype).cuda(), g), g)   # keep the perturbed field solenoidal
        eps = 1e-3 * scale / float(torch.sqrt(torch.mean(d.real ** 2 + d.imag ** 2)))
        cd = (J(u0 + eps * d) - J(u0 - eps * d)) / (2 * eps)
        cd2 = (J(u0 + 0.5 * eps * d) - J(u0 - 0.5 * eps * d)) / eps
        rich = (4 * cd2 - cd) / 3
        an = float(torch.sum(grad.real.double() * d.real.double() + grad.imag.double() * d.imag.double()))
        err_cd.append(abs(an - cd) / abs(cd))
        err_rich.append(abs(an - rich) / abs(rich))
    print(f"adjoint vs central differences ({dtype}, {N}^3, {steps} steps, 16 directions): "
          f"max rel err {max(err_cd):.2e} (eps), {max(err_rich):.2e} (Richardson)")
    assert max(err_cd) < 5e-5
    assert max(err_rich) < 1e-8


# --- gradients with respect to nu, A and eps (NSParams' leaves) --------------------------------

PARAM_FORCINGS = [
    (O.NoForcing(), lambda: h3.NoForcing()),
    (O.LinearForcing(A=0.15), lambda: h3.LinearForcing(A=0.15)),
    (O.BandForcing(eps=0.3, kmin=1.0, kmax=3.0), lambda: h3.BandForcing(eps=0.3, kmin=1.0, kmax=3.0)),
    (O.LundgrenForcing(nu=0.02), lambda: h3.LundgrenForcing(nu=0.02)),
]


@pytest.mark.parametrize("fi", range(len(PARAM_FORCINGS)))
def test_parameter_gradients_match_oracle(fi):
    """[d/dnu, d/dA, d/deps] accumulated on the GPU (FP64 sums of the stage
    cotangents) against the oracle's (pinned by central differences), for one
    RHS on a rough input and for a checkpointed 7-step rollout."""
    og, g = O.make_grid(16), h3.make_grid(16, dtype=torch.float64)
    u = O.random_field(7, og, O.config_spectrum)
    rough = u + 1e-2 * np.abs(u).max() * _rand_c(8, u.shape)
    ub = _rand_c(9, u.shape)
    of, gf = PARAM_FORCINGS[fi]
    op, gp = O.NSParams(nu=0.02, forcing=of), h3.NSParams(nu=0.02, forcing=gf())
    ref = O.rhs_param_vjp(rough, ub, op, og)
    pg = h3.adjoint.param_grad_buffer("cuda")
    h3.make_ns_rhs_vjp(g)(dev(rough), dev(ub), gp, 0.0, param_grad=pg)
    assert np.allclose(pg.cpu().numpy(), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    ref = np.zeros(3)
    O.rollout_vjp(O.make_ns_rhs(og), O.make_ns_rhs_vjp(og), u, ub, op, 2e-3, 7, checkpoint=3, pgrad=ref)
    pg.zero_()
    h3.rollout_vjp(h3.make_ns_rhs(g), dev(u), dev(ub), gp, 2e-3, 7, checkpoint=3, param_grad=pg)
    assert np.allclose(pg.cpu().numpy(), ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())


def test_grad_with_respect_to_params_through_autograd():
    """The reference's tests/test_grad.py:29-74 structure on the GPU: with nu
    and the forcing coefficient as tensors, torch.autograd through
    differentiable_rollout reaches them (finite, non-zero), and the values
    equal the oracle's."""
    og, g = O.make_grid(8), h3.make_grid(8, dtype=torch.float64)
    u = O.random_field(2, og)
    rhs = h3.make_ns_rhs(g)
    for of, make in ((O.LinearForcing(A=0.1), lambda c: h3.LinearForcing(A=c)),
                     (O.BandForcing(eps=0.1, kmin=1.0, kmax=2.0),
                      lambda c: h3.BandForcing(eps=c, kmin=1.0, kmax=2.0))):
        nu = torch.tensor(0.02, dtype=torch.float64, requires_grad=True)
        coef = torch.tensor(0.1, dtype=torch.float64, requires_grad=True)
        uN = h3.differentiable_rollout(rhs, dev(u), h3.NSParams(nu=nu, forcing=make(coef)), 1e-3, 2)
        (0.5 * torch.sum(uN.abs() ** 2)).backward()
        assert torch.isfinite(nu.grad) and nu.grad != 0 and torch.isfinite(coef.grad) and coef.grad != 0
        op = O.NSParams(nu=0.02, forcing=of)
        uNo = O.rollout(O.make_ns_rhs(og), u, op, 1e-3, 2)
        ref = np.zeros(3)
        O.rollout_vjp(O.make_ns_rhs(og), O.make_ns_rhs_vjp(og), u, uNo, op, 1e-3, 2, pgrad=ref)
        slot = 1 if isinstance(of, O.LinearForcing) else 2
        assert abs(float(nu.grad) - ref[0]) < 1e-11 * abs(ref[0])
        assert abs(float(coef.grad) - ref[slot]) < 1e-11 * abs(ref[slot])


def test_adjoint_config_parameter_gradients_central_differences():
    """The adjoint configuration (128^3, 200 steps, band forcing): dJ/dnu and
    dJ/deps from the adjoint against Richardson-extrapolated central
    differences of J in nu and in eps (float64)."""
    g = h3.make_grid(128, dtype=torch.float64)
    rhs = h3.make_ns_rhs(g)
    u0 = h3.random_field(0, g, h3.config_spectrum)
    W = _shell_energy_weights(g, 8, 16)
    nu0, eps0, steps = 1 / 600, 0.1, 200

    def params(nu, eps):
        return h3.NSParams(nu=nu, forcing=h3.BandForcing(eps=eps, kmin=1.0, kmax=2.0))

    def J(nu, eps):
        uN = h3.rollout(rhs, u0, params(nu, eps), 2e-3, steps)
        return float(torch.sum(W * (uN.real ** 2 + uN.imag ** 2)))
    nu = torch.tensor(nu0, dtype=torch.float64, requires_grad=True)
    eps = torch.tensor(eps0, dtype=torch.float64, requires_grad=True)
    uN = h3.differentiable_rollout(rhs, u0, params(nu, eps), 2e-3, steps, checkpoint=20)
    torch.sum(W * (uN.real ** 2 + uN.imag ** 2)).backward()
    for name, an, f, x0 in (("nu", float(nu.grad), lambda h: J(nu0 + h, eps0), nu0),
                            ("eps", float(eps.grad), lambda h: J(nu0, eps0 + h), eps0)):
        h = 1e-3 * x0
        cd = (f(h) - f(-h)) / (2 * h)
        cd2 = (f(0.5 * h) - f(-0.5 * h)) / h
        rich = (4 * cd2 - cd) / 3
        print(f"dJ/d{name}: adjoint {an:.10e}, Richardson central difference {rich:.10e}")
        assert abs(an - rich) < 1e-8 * abs(rich)
