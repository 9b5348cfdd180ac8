"""Geometry and spectral operators on the MI355X (grid.py of hit3d).

The state is one complex CUDA/HIP tensor of shape ``(3, Nx, Ny, Nz//2+1)``
(components first, rfft-halved axis last), complex64 for a float32 grid and
complex128 for a float64 grid. Every operator here runs a hand-written HIP
kernel through the C ABI (include/hit3d.h); torch only provides the device
memory and the current stream.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib

__all__ = [
    "Grid", "make_grid", "spectral_shape", "ddx", "ddy", "ddz", "laplacian",
    "dealias", "project", "to_physical", "to_spectral",
]

_DTYPES = {torch.float32: 4, torch.float64: 8}
_COMPLEX = {torch.float32: torch.complex64, torch.float64: torch.complex128}


def _as_torch_dtype(dtype):
    if isinstance(dtype, torch.dtype):
        return dtype
    d = np.dtype(dtype)
    if d == np.float32:
        return torch.float32
    if d == np.float64:
        return torch.float64
    raise ValueError(f"unsupported grid dtype {dtype}")


@dataclass(frozen=True, eq=False)
class Grid:
    """Box parameters plus lazily built per-mode tables (grid.py:41-86).

    ``kx, ky, kz`` are broadcast-shaped ``(Nx,1,1)``, ``(1,Ny,1)``,
    ``(1,1,Nz//2+1)``; ``k2``, ``inv_k2``, ``dealias`` (bool) live on the
    spectral grid and are produced by the same device code the kernels use;
    ``weights`` ``(1,1,Nz//2+1)`` are the Hermitian shell-sum weights.
    """

    Nx: int
    Ny: int
    Nz: int
    Lx: float
    Ly: float
    Lz: float
    dtype: torch.dtype
    device: torch.device
    _cache: dict = field(default_factory=dict, repr=False, compare=False)

    def __repr__(self):
        name = "float32" if self.dtype == torch.float32 else "float64"
        return (f"Grid[{name}]({self.Nx}x{self.Ny}x{self.Nz}, "
                f"L=({self.Lx:g}, {self.Ly:g}, {self.Lz:g}))")

    # -- C ABI plumbing ------------------------------------------------------
    @property
    def desc(self):
        d = self._cache.get("desc")
        if d is None:
            d = _lib.GridDesc(self.Nx, self.Ny, self.Nz, self.Lx, self.Ly, self.Lz,
                              _DTYPES[self.dtype], 0)
            self._cache["desc"] = d
        return ctypes.byref(d)

    @property
    def complex_dtype(self):
        return _COMPLEX[self.dtype]

    def workspace(self, op, ncomp=3, dealiased=1):
        """Device scratch for `op`, allocated once per (op, ncomp, layout)."""
        key = ("ws", op, ncomp, dealiased)
        ws = self._cache.get(key)
        if ws is None:
            n = ctypes.c_size_t()
            _lib.call("hit3d_workspace_size", self.desc, op, ncomp, dealiased, ctypes.byref(n))
            ws = torch.empty(max(int(n.value), 256), dtype=torch.uint8, device=self.device)
            self._cache[key] = ws
        return ws

    def _tables(self):
        t = self._cache.get("tables")
        if t is None:
            shape = (self.Nx, self.Ny, self.Nz // 2 + 1)
            k2 = torch.empty(shape, dtype=self.dtype, device=self.device)
            ik2 = torch.empty_like(k2)
            mask = torch.empty(shape, dtype=torch.uint8, device=self.device)
            shell = torch.empty(shape, dtype=torch.int32, device=self.device)
            _lib.call("hit3d_grid_tables", self.desc, k2.data_ptr(), ik2.data_ptr(),
                      mask.data_ptr(), shell.data_ptr(), stream())
            t = {"k2": k2, "inv_k2": ik2, "dealias": mask.bool(), "shell": shell}
            self._cache["tables"] = t
        return t

    def _axis(self, name):
        t = self._cache.get(name)
        if t is None:
            N, L, shape = {"kx": (self.Nx, self.Lx, (-1, 1, 1)),
                           "ky": (self.Ny, self.Ly, (1, -1, 1)),
                           "kz": (self.Nz, self.Lz, (1, 1, -1))}[name]
            n = np.arange(N // 2 + 1, dtype=float) if name == "kz" else np.fft.fftfreq(N, 1 / N)
            npd = np.float32 if self.dtype == torch.float32 else np.float64
            t = torch.from_numpy((2 * np.pi / L * n).astype(npd).reshape(shape)).to(self.device)
            self._cache[name] = t
        return t

    @property
    def kx(self):
        return self._axis("kx")

    @property
    def ky(self):
        return self._axis("ky")

    @property
    def kz(self):
        return self._axis("kz")

    @property
    def k2(self):
        return self._tables()["k2"]

    @property
    def inv_k2(self):
        return self._tables()["inv_k2"]

    @property
    def dealias(self):
        return self._tables()["dealias"]

    @property
    def shell(self):
        return self._tables()["shell"]

    @property
    def weights(self):
        t = self._cache.get("weights")
        if t is None:
            w = np.full(self.Nz // 2 + 1, 2.0)
            w[0] = 1
            if self.Nz % 2 == 0:
                w[-1] = 1
            t = torch.tensor(w.reshape(1, 1, -1), dtype=self.dtype, device=self.device)
            self._cache["weights"] = t
        return t


def stream():
    """The current torch stream as the C ABI's `void* stream`."""
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def make_grid(Nx, Ny=None, Nz=None, *, Lx=2 * np.pi, Ly=2 * np.pi, Lz=2 * np.pi,
              dtype=torch.float32, device="cuda"):
    """grid.py:89-132. Sizes must be powers of two in [4, 1024]."""
    Ny = Nx if Ny is None else Ny
    Nz = Nx if Nz is None else Nz
    for n in (Nx, Ny, Nz):
        if n < 4 or n > 1024 or (n & (n - 1)):
            raise ValueError(f"grid sizes must be powers of two in [4, 1024], got {n}")
    dev = torch.device(device)
    if dev.type != "cuda":
        raise ValueError("hit3d runs on the GPU: device must be a cuda (HIP) device")
    return Grid(int(Nx), int(Ny), int(Nz), float(Lx), float(Ly), float(Lz),
                _as_torch_dtype(dtype), dev)


def spectral_shape(grid):
    """grid.py:135-137."""
    return (3, grid.Nx, grid.Ny, grid.Nz // 2 + 1)


def check_spectral(u_hat, grid, ncomp=None, name="u_hat"):
    if not isinstance(u_hat, torch.Tensor) or not u_hat.is_cuda:
        raise ValueError(f"{name} must be a CUDA (HIP) tensor")
    if u_hat.dtype != grid.complex_dtype:
        raise ValueError(f"{name} must be {grid.complex_dtype}, got {u_hat.dtype}")
    tail = (grid.Nx, grid.Ny, grid.Nz // 2 + 1)
    if u_hat.dim() != 4 or tuple(u_hat.shape[1:]) != tail or (
            ncomp is not None and u_hat.shape[0] != ncomp):
        want = (ncomp if ncomp is not None else "ncomp",) + tail
        raise ValueError(f"{name} must have shape {want}, got {tuple(u_hat.shape)}")
    return u_hat.contiguous()


# Python attribute that records "no content outside the 2/3 box" on a state
# tensor, together with the tensor's version counter at that moment (any
# in-place write bumps the counter and voids the mark).
DEALIASED_MARK = "_hit3d_dealiased_version"


def mark_dealiased(u_hat):
    """Records on the tensor that it has no content outside the 2/3 box. The
    library's own producers of such states (the pruned step, the RHS on a
    dealiased state, random_field, dealias) mark their outputs, so that a
    chain rollout -> rollout never reads the device to decide the layout."""
    setattr(u_hat, DEALIASED_MARK, u_hat._version)
    return u_hat


def _new_like(u):
    return torch.empty_like(u, memory_format=torch.contiguous_format)


def dealias(u_hat, grid):
    """grid.py:166-168."""
    u = check_spectral(u_hat, grid)
    out = _new_like(u)
    _lib.call("hit3d_dealias_project", grid.desc, u.data_ptr(), out.data_ptr(), u.shape[0], 1, 0,
              stream())
    return mark_dealiased(out)


def project(u_hat, grid):
    """grid.py:171-178 (identity at k = 0, where inv_k2 = 0)."""
    u = check_spectral(u_hat, grid, 3)
    out = _new_like(u)
    _lib.call("hit3d_dealias_project", grid.desc, u.data_ptr(), out.data_ptr(), 3, 0, 1, stream())
    return out


def _derivative(u_hat, grid, op):
    squeeze = u_hat.dim() == 3
    u = check_spectral(u_hat.unsqueeze(0) if squeeze else u_hat, grid)
    out = _new_like(u)
    _lib.call("hit3d_spectral_derivative", grid.desc, u.data_ptr(), out.data_ptr(), u.shape[0], op,
              stream())
    return out[0] if squeeze else out


def ddx(u_hat, grid):
    """grid.py:146-148: i kx u."""
    return _derivative(u_hat, grid, 0)


def ddy(u_hat, grid):
    """grid.py:151-153."""
    return _derivative(u_hat, grid, 1)


def ddz(u_hat, grid):
    """grid.py:156-158."""
    return _derivative(u_hat, grid, 2)


def laplacian(u_hat, grid):
    """grid.py:161-163: -k2 u."""
    return _derivative(u_hat, grid, 3)


def to_physical(u_hat, grid):
    """grid.py:181-183: irfftn over the last three axes."""
    squeeze = u_hat.dim() == 3
    u = check_spectral(u_hat.unsqueeze(0) if squeeze else u_hat, grid)
    nc = u.shape[0]
    out = torch.empty((nc, grid.Nx, grid.Ny, grid.Nz), dtype=grid.dtype, device=u.device)
    ws = grid.workspace(_lib.OP_TO_PHYSICAL, nc)
    _lib.call("hit3d_to_physical", grid.desc, u.data_ptr(), out.data_ptr(), nc, ws.data_ptr(),
              ws.numel(), stream())
    return out[0] if squeeze else out


def to_spectral(u, grid):
    """grid.py:186-188: rfftn over the last three axes."""
    squeeze = u.dim() == 3
    x = u.unsqueeze(0) if squeeze else u
    if not isinstance(x, torch.Tensor) or not x.is_cuda or x.dtype != grid.dtype:
        raise ValueError(f"u must be a CUDA tensor of dtype {grid.dtype}")
    if x.dim() != 4 or tuple(x.shape[1:]) != (grid.Nx, grid.Ny, grid.Nz):
        raise ValueError(f"u must have shape (ncomp, {grid.Nx}, {grid.Ny}, {grid.Nz})")
    x = x.contiguous()
    nc = x.shape[0]
    out = torch.empty((nc, grid.Nx, grid.Ny, grid.Nz // 2 + 1), dtype=grid.complex_dtype,
                      device=x.device)
    ws = grid.workspace(_lib.OP_TO_SPECTRAL, nc)
    _lib.call("hit3d_to_spectral", grid.desc, x.data_ptr(), out.data_ptr(), nc, ws.data_ptr(),
              ws.numel(), stream())
    return out[0] if squeeze else out
