"""Learned subgrid closure: a small periodic 3-D conv net run by the BF16
MFMA kernels of the library (hit3d_conv_* in include/hit3d.h).

The net maps the filtered velocity-gradient tensor (9 channels, physical
space, scaled by `in_scale`) to the subgrid stress tau_ij (6 channels, order
00, 01, 02, 11, 12, 22, scaled back by `out_scale`) through three 3x3x3
convolutions (9 -> C, C -> C, C -> 6) with GELU between them. Parameters
are FP32 masters with BF16 packed copies for the MFMA operands; gradients
come from the library's backward kernels and Adam runs on the device.
There is no reference counterpart: the closure is an addition to hit3d.
"""

from __future__ import annotations

import ctypes
import math

import torch

from . import _lib
from .grid import stream

__all__ = ["ConvLayer", "ClosureNet"]


def _pad(c):
    return 16 if c <= 16 else 32


class ConvLayer:
    """One periodic 3x3x3 convolution on a (d, h, w) voxel grid.

    w: fp32 [cout][27][cin] (offset o = 9*(da+1) + 3*(db+1) + (dc+1)),
    b: fp32 [cout]. `gelu` selects the GELU output (pre-activation kept for
    the backward pass) or a plain fp32 output."""

    def __init__(self, dhw, cin, cout, gelu, device, generator=None):
        if cin > 32 or cout > 32:
            raise ValueError("at most 32 channels per layer")
        d, h, w = dhw
        self.dhw, self.cin, self.cout, self.gelu = dhw, cin, cout, bool(gelu)
        self.desc = _lib.ConvDesc(d, h, w, cin, cout, _pad(cin), _pad(cout))
        kp, kt, nb = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_size_t()
        _lib.call("hit3d_conv_sizes", ctypes.byref(self.desc), ctypes.byref(kp), ctypes.byref(kt),
                  ctypes.byref(nb))
        self.V = d * h * w
        std = math.sqrt((2.0 if gelu else 1.0) / (27 * cin))
        self.w = (torch.randn(cout, 27, cin, generator=generator) * std).to(device)
        self.b = torch.zeros(cout, device=device)
        self.wp = torch.empty(_pad(cout), kp.value, dtype=torch.bfloat16, device=device)
        self.wt = torch.empty(_pad(cin), kt.value, dtype=torch.bfloat16, device=device)
        self.work = torch.empty(max(nb.value, 256), dtype=torch.uint8, device=device)
        self.gw, self.gb = torch.zeros_like(self.w), torch.zeros_like(self.b)
        self._gw, self._gb = torch.empty_like(self.w), torch.empty_like(self.b)
        self.mw, self.vw = torch.zeros_like(self.w), torch.zeros_like(self.w)
        self.mb, self.vb = torch.zeros_like(self.b), torch.zeros_like(self.b)
        self.pack()

    @property
    def d(self):
        return ctypes.byref(self.desc)

    def pack(self):
        _lib.call("hit3d_conv_pack", self.d, self.w.data_ptr(), self.wp.data_ptr(),
                  self.wt.data_ptr(), stream())

    def forward(self, x, keep_pre=True):
        """x bf16 [V][cin_pad] -> (pre fp32 [V][cout_pad], act bf16 [V][cout_pad])
        with GELU (pre None unless keep_pre: only a backward pass reads it),
        else out fp32 [V][cout]."""
        dev = x.device
        if self.gelu:
            pre = torch.empty(self.V, self.desc.cout_pad, device=dev) if keep_pre else None
            act = torch.empty(self.V, self.desc.cout_pad, dtype=torch.bfloat16, device=dev)
            _lib.call("hit3d_conv_forward", self.d, x.data_ptr(), self.wp.data_ptr(),
                      self.b.data_ptr(), 1, None if pre is None else pre.data_ptr(), act.data_ptr(),
                      None, stream())
            return pre, act
        out = torch.empty(self.V, self.cout, device=dev)
        _lib.call("hit3d_conv_forward", self.d, x.data_ptr(), self.wp.data_ptr(), self.b.data_ptr(),
                  0, None, None, out.data_ptr(), stream())
        return out

    def backward(self, x, pre, grad_out, need_grad_x=True):
        """Adds this call's weight and bias gradients to (gw, gb); returns
        grad_x fp32 [V][cin_pad] (or None)."""
        gx = torch.empty(self.V, self.desc.cin_pad, device=x.device) if need_grad_x else None
        _lib.call("hit3d_conv_backward", self.d, x.data_ptr(), self.wt.data_ptr(),
                  None if pre is None else pre.data_ptr(), grad_out.contiguous().data_ptr(),
                  1 if self.gelu else 0, None if gx is None else gx.data_ptr(),
                  self._gw.data_ptr(), self._gb.data_ptr(), self.work.data_ptr(), self.work.numel(),
                  stream())
        self.gw += self._gw
        self.gb += self._gb
        return gx

    def zero_grad(self):
        self.gw.zero_()
        self.gb.zero_()

    def adam(self, lr, t, beta1=0.9, beta2=0.999, eps=1e-8):
        for p, g, m, v in ((self.w, self.gw, self.mw, self.vw), (self.b, self.gb, self.mb, self.vb)):
            _lib.call("hit3d_adam", p.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(), p.numel(),
                      float(lr), float(beta1), float(beta2), float(eps), int(t), stream())
        self.pack()

    def state(self):
        return {"w": self.w.clone(), "b": self.b.clone()}

    def load(self, st):
        self.w.copy_(st["w"])
        self.b.copy_(st["b"])
        self.pack()


class ClosureNet:
    """tau = out_scale * net(in_scale * g): g fp32 [9][V] -> tau fp32 [6][V]
    on the LES grid (V = Nx*Ny*Nz voxels, physical layout), with tau set to
    0 where it would transfer energy to the resolved scales (-tau_ij S_ij <
    0) while `clip_backscatter` is on (hit3d_stress_unpack_clip)."""

    def __init__(self, grid, channels=32, seed=0, in_scale=None, out_scale=None):
        dev = grid.device
        self.grid = grid
        self.dhw = (grid.Nx, grid.Ny, grid.Nz)
        self.V = grid.Nx * grid.Ny * grid.Nz
        gen = torch.Generator().manual_seed(int(seed))
        self.layers = [ConvLayer(self.dhw, 9, channels, True, dev, gen),
                       ConvLayer(self.dhw, channels, channels, True, dev, gen),
                       ConvLayer(self.dhw, channels, 6, False, dev, gen)]
        self.in_scale = torch.ones(9, device=dev) if in_scale is None else in_scale.float().to(dev)
        self.out_scale = torch.ones(6, device=dev) if out_scale is None else out_scale.float().to(dev)
        self.t = 0
        # stands for the parameters in torch's graph: the net's output depends
        # on it, so losses reach the net's backward even when the input state
        # does not require grad (the parameters themselves live in the layers)
        self.token = torch.zeros(1, device=dev, requires_grad=True)
        # scalar gain on the closure force, trained a posteriori; backscatter
        # clipping (tau set to 0 where -tau_ij S_ij < 0) keeps the modelled
        # subgrid transfer dissipative
        self.gain = torch.ones(1, device=dev, requires_grad=True)
        self.clip_backscatter = True
        # spectral shape of the closure force, exp(tilt (|k|/kc - 1/2)) per
        # mode, kc = N/3; a device tensor updated in place, so captured LES
        # steps see the current tilt
        self.tilt = 0.0
        self._shape = None

    def spectral_shape(self):
        """exp(tilt (|k|/kc - 1/2)) on the net's grid (real, [Nx][Ny][Nz/2+1]),
        |k| and the cut kc both physical wavenumbers."""
        if self._shape is None:
            self._shape = torch.empty(self.grid.k2.shape, dtype=self.grid.dtype, device=self.grid.device)
            self._fill_shape()
        return self._shape

    def _fill_shape(self):
        # kc: the 2/3 cut as a wavenumber, (2 pi / L_i) (N_i // 3), smallest over the axes
        g = self.grid
        kc = min(2 * math.pi / L * (N // 3) for N, L in ((g.Nx, g.Lx), (g.Ny, g.Ly), (g.Nz, g.Lz)))
        with torch.no_grad():
            self._shape.copy_(torch.exp(self.tilt * (torch.sqrt(self.grid.k2) / kc - 0.5)))

    def set_tilt(self, tilt):
        self.tilt = float(tilt)
        if self._shape is not None:
            self._fill_shape()

    def set_scales(self, in_scale, out_scale):
        self.in_scale = in_scale.float().to(self.in_scale.device).contiguous()
        self.out_scale = out_scale.float().to(self.out_scale.device).contiguous()

    def forward_saved(self, g, keep=True):
        """(tau fp32 [6][V], saved tensors for `backward_saved`); keep=False
        skips the pre-activations that only the backward pass reads."""
        g = g.contiguous()
        x0 = torch.empty(self.V, 16, dtype=torch.bfloat16, device=g.device)
        _lib.call("hit3d_channels_pack", g.data_ptr(), 9, self.V, self.in_scale.data_ptr(), 16,
                  x0.data_ptr(), stream())
        p1, a1 = self.layers[0].forward(x0, keep)
        p2, a2 = self.layers[1].forward(a1, keep)
        out = self.layers[2].forward(a2)
        tau = torch.empty(6, self.V, device=g.device)
        keep = None
        if self.clip_backscatter:
            keep = torch.empty(self.V, dtype=torch.uint8, device=g.device)
            _lib.call("hit3d_stress_unpack_clip", out.data_ptr(), self.V, self.out_scale.data_ptr(),
                      g.data_ptr(), tau.data_ptr(), keep.data_ptr(), stream())
        else:
            _lib.call("hit3d_channels_unpack", out.data_ptr(), 6, 6, self.V, self.out_scale.data_ptr(),
                      tau.data_ptr(), stream())
        return tau, (x0, p1, a1, p2, a2, keep)

    def __call__(self, g):
        """tau for inference (nothing kept for a backward pass)."""
        return self.forward_saved(g, keep=False)[0]

    def backward_saved(self, saved, grad_tau, need_grad_g=True):
        """Accumulates the parameter gradients; returns d/dg (fp32 [9][V])."""
        x0, p1, a1, p2, a2, keep = saved
        if keep is not None:   # clipped entries are constant (zero): no gradient flows back
            grad_tau = grad_tau * keep.to(grad_tau.dtype)[None, :]
        gout = (grad_tau * self.out_scale[:, None]).t().contiguous()      # [V][6]
        ga2 = self.layers[2].backward(a2, None, gout)
        ga1 = self.layers[1].backward(a1, p2, ga2)
        gx0 = self.layers[0].backward(x0, p1, ga1, need_grad_x=need_grad_g)
        if not need_grad_g:
            return None
        gg = torch.empty(9, self.V, device=grad_tau.device)
        _lib.call("hit3d_channels_unpack", gx0.data_ptr(), 9, 16, self.V, self.in_scale.data_ptr(),
                  gg.data_ptr(), stream())
        return gg

    def zero_grad(self):
        for L in self.layers:
            L.zero_grad()

    def reset_optimizer(self):
        """Zero the Adam moments and step count (a new training phase)."""
        self.t = 0
        for L in self.layers:
            for m in (L.mw, L.vw, L.mb, L.vb):
                m.zero_()

    def adam_step(self, lr):
        self.t += 1
        for L in self.layers:
            L.adam(lr, self.t)

    def num_parameters(self):
        return sum(L.w.numel() + L.b.numel() for L in self.layers)

    def state(self):
        return {"layers": [L.state() for L in self.layers], "in_scale": self.in_scale.clone(),
                "out_scale": self.out_scale.clone(), "gain": self.gain.detach().clone(),
                "clip_backscatter": self.clip_backscatter, "tilt": self.tilt}

    def load(self, st):
        for L, s in zip(self.layers, st["layers"]):
            L.load(s)
        self.set_scales(st["in_scale"], st["out_scale"])
        with torch.no_grad():
            self.gain.copy_(st["gain"])
        self.clip_backscatter = st["clip_backscatter"]
        self.set_tilt(st.get("tilt", 0.0))


class _NetFn(torch.autograd.Function):
    """tau = net(g) with the library's backward kernels; parameter gradients
    accumulate in the net's layers (ClosureNet.zero_grad / adam_step)."""

    @staticmethod
    def forward(ctx, g, token, net):
        tau, saved = net.forward_saved(g)
        ctx.net, ctx.saved = net, saved
        return tau

    @staticmethod
    def backward(ctx, grad_tau):
        gg = ctx.net.backward_saved(ctx.saved, grad_tau.contiguous(), need_grad_g=ctx.needs_input_grad[0])
        ctx.saved = None
        return gg, torch.zeros_like(ctx.net.token), None


def apply_net(net, g):
    """Differentiable tau = net(g) (autograd through the library kernels).
    Outside torch.no_grad() the result always carries a graph, since the
    parameters stand behind `net.token`."""
    return _NetFn.apply(g, net.token, net)
