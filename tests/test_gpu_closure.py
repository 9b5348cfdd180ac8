"""The closure's BF16-MFMA convolutions and the filtered-DNS kernels on the
GPU: each layer against torch's float32 conv3d with circular padding on the
same bf16-rounded operands (the library accumulates in fp32, so forward and
plain backward agree to fp32 summation order), the GELU backward (whose
bf16 rounding of dL/dpre may differ by one bf16 ulp), Adam, the layout
packers, and the spectral truncation / products / stress divergence
against NumPy."""

import numpy as np
import pytest
import torch
import torch.nn.functional as F

import hit3d_oracle as O

pytestmark = pytest.mark.gpu

h3 = pytest.importorskip("hit3d")
from hit3d import _lib
from hit3d.closure import ClosureNet, ConvLayer, apply_net
from hit3d.grid import stream

DHW = (16, 16, 16)
V = 16 * 16 * 16


def torch_conv(x, w, b, dhw=DHW):
    """x [V][cin] fp32 -> [V][cout]: periodic 3x3x3 cross-correlation."""
    d, h, wd = dhw
    cout, _, cin = w.shape
    X = x[:, :cin].t().reshape(1, cin, d, h, wd)
    W = w.reshape(cout, 3, 3, 3, cin).permute(0, 4, 1, 2, 3)
    Y = F.conv3d(F.pad(X, (1, 1, 1, 1, 1, 1), mode="circular"), W, b)
    return Y.reshape(cout, -1).t()


def bf(x):
    return x.to(torch.bfloat16).float()


def rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


def make_input(cin, cpad, seed):
    g = torch.Generator().manual_seed(seed)
    x = torch.zeros(V, cpad)
    x[:, :cin] = torch.randn(V, cin, generator=g)
    return x.to(torch.bfloat16).cuda()


@pytest.mark.parametrize("cin,cout,gelu", [(9, 32, True), (32, 32, True), (32, 6, False), (16, 16, True)])
def test_conv_forward_matches_torch(cin, cout, gelu):
    L = ConvLayer(DHW, cin, cout, gelu, "cuda", torch.Generator().manual_seed(cin + cout))
    L.b.copy_(torch.linspace(-0.5, 0.5, cout))
    x = make_input(cin, L.desc.cin_pad, 1)
    ref = torch_conv(x.float(), bf(L.w), L.b)
    if gelu:
        pre, act = L.forward(x)
        assert rel(pre[:, :cout], ref) < 1e-5
        if cout < pre.shape[1]:
            assert float(pre[:, cout:].abs().max()) == 0.0   # zero-padded weights and bias
        gref = F.gelu(ref)
        assert float(((act[:, :cout].float() - gref).abs() - 2 ** -8 * gref.abs()).max()) <= 1e-6
        if cout < act.shape[1]:
            assert float(act[:, cout:].float().abs().max()) == 0.0
    else:
        out = L.forward(x)
        assert out.shape == (V, cout)
        assert rel(out, ref) < 1e-5


@pytest.mark.parametrize("dhw", [(8, 16, 32), (4, 8, 64), (2, 32, 16), (16, 12, 32)])
@pytest.mark.parametrize("cin,cout", [(9, 32), (32, 32), (32, 6)])
def test_conv_forward_tiled_and_fallback_shapes(dhw, cin, cout):
    """The LDS-tiled forward (D % 4 == 0, H % 8 == 0: 4 x 8 x 16 voxel tiles
    with their periodic halo) and the per-wave fallback (the other shapes)
    against torch on anisotropic grids, where a transposed tile axis or a
    halo wrapped on the wrong axis would show."""
    d, h, wd = dhw
    v = d * h * wd
    L = ConvLayer(dhw, cin, cout, cout == 32, "cuda", torch.Generator().manual_seed(7 * cin + cout))
    L.b.copy_(torch.linspace(-0.5, 0.5, cout))
    g = torch.Generator().manual_seed(3)
    x = torch.zeros(v, L.desc.cin_pad)
    x[:, :cin] = torch.randn(v, cin, generator=g)
    x = x.to(torch.bfloat16).cuda()
    ref = torch_conv(x.float(), bf(L.w), L.b, dhw)
    got = L.forward(x)[0][:, :cout] if cout == 32 else L.forward(x)
    assert rel(got, ref) < 1e-5
    if cout == 32:   # inference form: no pre-activation kept, the same activations
        pre, act = L.forward(x)
        none, act2 = L.forward(x, keep_pre=False)
        assert none is None and torch.equal(act, act2)


def _torch_grads(x, w, b, gout_pre):
    """Gradients of sum(conv(x) * gout_pre) w.r.t. x, w, b in float32."""
    xr = x.clone().requires_grad_(True)
    wr = w.clone().requires_grad_(True)
    br = b.clone().requires_grad_(True)
    y = torch_conv(xr, wr, br)
    (y * gout_pre).sum().backward()
    return xr.grad, wr.grad, br.grad


def test_conv_backward_plain_matches_torch():
    L = ConvLayer(DHW, 32, 6, False, "cuda", torch.Generator().manual_seed(5))
    x = make_input(32, 32, 2)
    G = torch.randn(V, 6, generator=torch.Generator().manual_seed(3)).cuda()
    L.zero_grad()
    gx = L.backward(x, None, G)
    rx, rw, rb = _torch_grads(x.float(), bf(L.w), L.b, bf(G))   # the kernel rounds G to bf16
    assert rel(gx[:, :32], rx) < 1e-5
    assert rel(L.gw, rw) < 1e-5
    assert rel(L.gb, rb) < 1e-5
    # gradients accumulate over calls until zero_grad
    L.backward(x, None, G, need_grad_x=False)
    assert rel(L.gw, 2 * rw) < 1e-5


@pytest.mark.parametrize("cin,cout", [(9, 32), (32, 32)])
def test_conv_backward_gelu_matches_torch(cin, cout):
    L = ConvLayer(DHW, cin, cout, True, "cuda", torch.Generator().manual_seed(7))
    x = make_input(cin, L.desc.cin_pad, 4)
    pre, _ = L.forward(x)
    G = torch.randn(V, L.desc.cout_pad, generator=torch.Generator().manual_seed(8)).cuda()
    L.zero_grad()
    gx = L.backward(x, pre, G)
    p = pre[:, :cout]
    dgelu = 0.5 * (1 + torch.erf(p / np.sqrt(2))) + p * torch.exp(-0.5 * p * p) / np.sqrt(2 * np.pi)
    dP = bf(G[:, :cout] * dgelu)
    rx, rw, rb = _torch_grads(x.float(), bf(L.w), L.b, dP)
    assert rel(gx[:, :cin], rx[:, :cin]) < 5e-3
    assert rel(L.gw, rw) < 5e-3
    assert rel(L.gb, rb) < 5e-3


def test_adam_matches_torch():
    n = 1000
    g = torch.Generator().manual_seed(9)
    p0, grads = torch.randn(n, generator=g), [torch.randn(n, generator=g) for _ in range(3)]
    p, m, v = p0.clone().cuda(), torch.zeros(n).cuda(), torch.zeros(n).cuda()
    ref = p0.clone().requires_grad_(True)
    opt = torch.optim.Adam([ref], lr=1e-2)
    for t, gr in enumerate(grads, 1):
        _lib.call("hit3d_adam", p.data_ptr(), gr.cuda().data_ptr(), m.data_ptr(), v.data_ptr(), n,
                  1e-2, 0.9, 0.999, 1e-8, t, stream())
        ref.grad = gr.clone()
        opt.step()
    assert float((p.cpu() - ref.detach()).abs().max()) < 1e-6


def test_net_autograd_matches_layerwise_torch():
    """d/dg of sum(tau * G) through apply_net against the same composition in
    torch float32 (bf16 rounding where the kernels round)."""
    grid = h3.make_grid(16, dtype=torch.float32)
    net = ClosureNet(grid, channels=32, seed=3)
    net.clip_backscatter = False          # the raw net against torch (clipping is tested in test_gpu_les)
    net.set_scales(torch.full((9,), 0.5), torch.full((6,), 2.0))
    gen = torch.Generator().manual_seed(11)
    g = torch.randn(9, V, generator=gen).cuda().requires_grad_(True)
    G = torch.randn(6, V, generator=gen).cuda()
    net.zero_grad()
    tau = apply_net(net, g)
    (tau * G).sum().backward()
    # torch reference of the forward
    L1, L2, L3 = net.layers
    x0 = bf(g.detach().t() * 0.5)
    p1 = torch_conv(x0, bf(L1.w), L1.b)
    p2 = torch_conv(bf(F.gelu(p1)), bf(L2.w), L2.b)
    out = torch_conv(bf(F.gelu(p2)), bf(L3.w), L3.b)
    assert rel(tau.detach(), (out * 2.0).t()) < 2e-2
    # and of the input gradient, in float32 without the bf16 roundings (loose)
    gr = g.detach().clone().requires_grad_(True)
    xr = gr.t() * 0.5
    y = torch_conv(F.gelu(torch_conv(F.gelu(torch_conv(xr, L1.w, L1.b)), L2.w, L2.b)), L3.w, L3.b)
    ((y * 2.0).t() * G).sum().backward()
    assert rel(g.grad, gr.grad) < 5e-2
    assert float(net.layers[0].gw.abs().max()) > 0


def test_channels_pack_unpack():
    gen = torch.Generator().manual_seed(12)
    soa = torch.randn(9, V, generator=gen).cuda()
    s = torch.linspace(0.5, 2.0, 9).cuda()
    aos = torch.empty(V, 16, dtype=torch.bfloat16, device="cuda")
    _lib.call("hit3d_channels_pack", soa.data_ptr(), 9, V, s.data_ptr(), 16, aos.data_ptr(), stream())
    assert torch.equal(aos[:, :9], (soa * s[:, None]).t().to(torch.bfloat16))
    assert float(aos[:, 9:].float().abs().max()) == 0.0
    f = torch.randn(V, 16, generator=gen).cuda()
    back = torch.empty(9, V, device="cuda")
    _lib.call("hit3d_channels_unpack", f.data_ptr(), 9, 16, V, s.data_ptr(), back.data_ptr(), stream())
    assert torch.equal(back, (f[:, :9] * s[None, :]).t())


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_spectral_truncate_matches_numpy(dtype):
    fine, coarse = h3.make_grid(32, dtype=dtype), h3.make_grid(16, dtype=dtype)
    og = O.make_grid(32, dtype=np.float64)
    u = O.random_field(5, og, O.config_spectrum)
    uc = torch.empty(3, 16, 16, 9, dtype=coarse.complex_dtype, device="cuda")
    ud = torch.from_numpy(u).to(fine.complex_dtype).cuda()
    _lib.call("hit3d_spectral_truncate", fine.desc, ud.data_ptr(), coarse.desc, uc.data_ptr(), 3, 5.0, stream())
    # NumPy: keep |n| <= 5 on the coarse index set, scale (16/32)^3
    n = np.fft.fftfreq(16, 1 / 16).astype(int)
    ref = np.zeros((3, 16, 16, 9), dtype=np.complex128)
    for ix, nx in enumerate(n):
        for iy, ny in enumerate(n):
            for iz in range(9):
                if abs(nx) < 8 and abs(ny) < 8 and iz < 8 and nx * nx + ny * ny + iz * iz <= 25:
                    ref[:, ix, iy, iz] = u[:, nx % 32, ny % 32, iz] / 8.0
    tol = 1e-6 if dtype == torch.float32 else 1e-15
    assert np.abs(uc.cpu().numpy() - ref).max() <= tol * np.abs(ref).max()
    # physical meaning: the filtered field on the coarse grid equals the fine
    # field's sharp-filtered values at the coarse points
    filt = u * (og.k2 <= 25.0)
    phys_f = O.to_physical(filt, og)[:, ::2, ::2, ::2]
    phys_c = O.to_physical(ref, O.make_grid(16, dtype=np.float64))
    assert np.abs(phys_f - phys_c).max() < 1e-12 * np.abs(phys_f).max()


def test_products_and_stress_divergence_match_numpy():
    grid = h3.make_grid(16, dtype=torch.float64)
    og = O.make_grid(16, dtype=np.float64)
    rng = np.random.default_rng(13)
    u = rng.standard_normal((3, 16, 16, 16))
    uu = torch.empty(6, 16, 16, 16, dtype=torch.float64, device="cuda")
    _lib.call("hit3d_velocity_products", grid.desc, torch.from_numpy(u).cuda().data_ptr(), uu.data_ptr(),
              stream())
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    assert np.array_equal(uu.cpu().numpy(), np.stack([u[i] * u[j] for i, j in pairs]))
    tau = O.to_spectral(rng.standard_normal((6, 16, 16, 16)), og)
    out = torch.empty(3, 16, 16, 9, dtype=torch.complex128, device="cuda")
    _lib.call("hit3d_stress_divergence", grid.desc, torch.from_numpy(tau).cuda().data_ptr(),
              out.data_ptr(), 0, stream())
    k = [og.kx, og.ky, og.kz]
    T = {}
    for s, (i, j) in enumerate(pairs):
        T[i, j] = T[j, i] = tau[s]
    ref = np.stack([-1j * sum(k[j] * T[i, j] for j in range(3)) for i in range(3)])
    assert np.abs(out.cpu().numpy() - ref).max() < 1e-13 * np.abs(ref).max()
    # adjoint identity: Re<g, D tau> = Re<D* g, tau> (sum over all stored entries)
    gF = rng.standard_normal((3, 16, 16, 9)) + 1j * rng.standard_normal((3, 16, 16, 9))
    adj = torch.empty(6, 16, 16, 9, dtype=torch.complex128, device="cuda")
    _lib.call("hit3d_stress_divergence", grid.desc, torch.from_numpy(gF).cuda().data_ptr(),
              adj.data_ptr(), 1, stream())
    lhs = np.real(np.sum(np.conj(gF) * ref))
    rhs = np.real(np.sum(np.conj(adj.cpu().numpy()) * tau))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
