"""hit3d on MI355X: the pure-function API of hit3d (nfrank22-gh/3DHIT,
src/hit3d/__init__.py:15-78) over a complex state (3, Nx, Ny, Nz//2+1), run by
hand-written gfx950 HIP kernels behind the C ABI of include/hit3d.h.

State tensors are torch CUDA (HIP) tensors: complex64 for float32 grids,
complex128 for float64 grids. There is no CPU path: importing works without
a GPU, but every operator needs the built libhit3d_hip.so and a device.
"""

from .adjoint import differentiable_rollout, make_ns_rhs_vjp, rk4_step_vjp, rollout_vjp
from .callbacks import FieldWriter, Hdf5Sink, MemorySink, save_series
from .diagnostics import (
    component_spectra, dissipation, energy, energy_budget, energy_budget_cb,
    energy_spectrum, enstrophy,
)
from .fields import (
    CONFIG_SPECTRUM_C, config_spectrum, default_spectrum, random_field, spectral_state,
    taylor_green, uniform_noise,
)
from .forcing import BandForcing, LinearForcing, LundgrenForcing, NoForcing
from .grid import (
    Grid, ddx, ddy, ddz, dealias, laplacian, make_grid, project, spectral_shape, to_physical,
    to_spectral,
)
from .integrators import Callback, Diagnostic, State, evolve, if_rk4_step, rk4_step, rollout
from .labels import label, run_label
from .navierstokes import NSParams, is_dealiased, linear_operator, make_ns_rhs
from .report import load_run, read_series

__all__ = [
    "Grid", "make_grid", "spectral_shape", "ddx", "ddy", "ddz", "laplacian",
    "dealias", "project", "to_physical", "to_spectral",
    "spectral_state", "random_field", "taylor_green", "default_spectrum", "config_spectrum",
    "uniform_noise", "CONFIG_SPECTRUM_C",
    "NoForcing", "LinearForcing", "BandForcing", "LundgrenForcing",
    "NSParams", "make_ns_rhs", "linear_operator", "is_dealiased",
    "rk4_step", "if_rk4_step", "rollout", "evolve", "State", "Callback", "Diagnostic",
    "FieldWriter", "save_series", "Hdf5Sink", "MemorySink", "load_run", "read_series",
    "label", "run_label",
    "energy", "enstrophy", "dissipation", "energy_spectrum", "component_spectra",
    "energy_budget", "energy_budget_cb",
    "make_ns_rhs_vjp", "rk4_step_vjp", "rollout_vjp", "differentiable_rollout",
]
