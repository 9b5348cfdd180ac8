"""ctypes binding of the C ABI in include/hit3d.h (libhit3d_hip.so, built
in-tree by 3dhit_amd/build.py). There is no fallback: if the library is
missing or a call fails, this module raises."""

from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhit3d_hip.so")

HIT3D_OK, HIT3D_ERR_INVALID, HIT3D_ERR_HIP, HIT3D_ERR_WORKSPACE = 0, 1, 2, 3
OP_TO_PHYSICAL, OP_TO_SPECTRAL, OP_RHS, OP_RK4, OP_SPECTRUM, OP_REDUCE, OP_RHS_VJP, OP_RK4_VJP = range(8)
FORCING_NONE, FORCING_LINEAR, FORCING_BAND, FORCING_LUNDGREN = range(4)


class GridDesc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int64), ("ny", ctypes.c_int64), ("nz", ctypes.c_int64),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("precision", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


class ForcingDesc(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("a", ctypes.c_double),
        ("eps", ctypes.c_double), ("kmin", ctypes.c_double), ("kmax", ctypes.c_double),
        ("nu", ctypes.c_double), ("energy", ctypes.c_double), ("tau", ctypes.c_double),
    ]


class ConvDesc(ctypes.Structure):
    _fields_ = [
        ("d", ctypes.c_int32), ("h", ctypes.c_int32), ("w", ctypes.c_int32),
        ("cin", ctypes.c_int32), ("cout", ctypes.c_int32),
        ("cin_pad", ctypes.c_int32), ("cout_pad", ctypes.c_int32),
    ]


_P = ctypes.c_void_p
_G = ctypes.POINTER(GridDesc)
_C = ctypes.POINTER(ConvDesc)
_F = ctypes.POINTER(ForcingDesc)
_SZ = ctypes.c_size_t
_I32, _I64, _D = ctypes.c_int32, ctypes.c_int64, ctypes.c_double

# name -> argtypes (restype is int for all)
SIGNATURES = {
    "hit3d_last_error": [ctypes.c_char_p, _SZ],
    "hit3d_version": [],
    "hit3d_workspace_size": [_G, _I32, _I64, _I32, ctypes.POINTER(_SZ)],
    "hit3d_shell_bins": [_G, ctypes.POINTER(_I64)],
    "hit3d_grid_tables": [_G, _P, _P, _P, _P, _P],
    "hit3d_to_physical": [_G, _P, _P, _I64, _P, _SZ, _P],
    "hit3d_to_spectral": [_G, _P, _P, _I64, _P, _SZ, _P],
    "hit3d_dealias_project": [_G, _P, _P, _I64, _I32, _I32, _P],
    "hit3d_spectral_derivative": [_G, _P, _P, _I64, _I32, _P],
    "hit3d_rhs": [_G, _P, _P, _D, _F, _D, _I32, _P, _SZ, _P],
    "hit3d_rk4_steps": [_G, _P, _D, _D, _I64, _D, _F, _I32, _P, _P, _SZ, _P],
    "hit3d_rk4_if_steps": [_G, _P, _D, _D, _I64, _D, _F, _I32, _P, _P, _SZ, _P],
    "hit3d_shell_spectrum": [_G, _P, _I32, _I32, _P, _P, _SZ, _P],
    "hit3d_energy_dissipation": [_G, _P, _D, _D, _D, _P, _P, _SZ, _P],
    "hit3d_forcing_term": [_G, _P, _F, _P, _P, _SZ, _P],
    "hit3d_energy_dissipation_noalloc": [_G, _P, _D, _D, _D, _P, _P],
    "hit3d_alias_energy": [_G, _P, _P, _P, _SZ, _P],
    "hit3d_uniform_noise": [_G, ctypes.c_uint64, _P, _P],
    "hit3d_shell_scale": [_G, _P, _I64, _P, _I64, _P],
    "hit3d_stream_copy": [_P, _P, _SZ, _I32, _P],
    "hit3d_profile_rk4": [_G, _P, _D, _D, _I64, _F, _I32, _P, _P, _P, _SZ, _P],
    "hit3d_rhs_vjp": [_G, _P, _P, _P, _D, _F, _D, _I32, _P, _P, _SZ, _P],
    "hit3d_rk4_step_vjp": [_G, _P, _P, _P, _D, _D, _D, _F, _I32, _P, _P, _P, _SZ, _P],
    "hit3d_rk4_steps_record": [_G, _P, _D, _D, _I64, _D, _F, _I32, _P, _P, _SZ, _P],
    "hit3d_conv_sizes": [_C, ctypes.POINTER(_I64), ctypes.POINTER(_I64), ctypes.POINTER(_SZ)],
    "hit3d_conv_pack": [_C, _P, _P, _P, _P],
    "hit3d_conv_forward": [_C, _P, _P, _P, _I32, _P, _P, _P, _P],
    "hit3d_conv_backward": [_C, _P, _P, _P, _P, _I32, _P, _P, _P, _P, _SZ, _P],
    "hit3d_adam": [_P, _P, _P, _P, _I64, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                   ctypes.c_float, _I64, _P],
    "hit3d_channels_pack": [_P, _I32, _I64, _P, _I32, _P, _P],
    "hit3d_channels_unpack": [_P, _I32, _I32, _I64, _P, _P, _P],
    "hit3d_stress_unpack_clip": [_P, _I64, _P, _P, _P, _P, _P],
    "hit3d_spectral_truncate": [_G, _P, _G, _P, _I64, _D, _P],
    "hit3d_velocity_products": [_G, _P, _P, _P],
    "hit3d_stress_divergence": [_G, _P, _P, _I32, _P],
    "hit3d_hold_energy": [_G, _P, _P, _P, _D, _D, _P, _P, _SZ, _P],
    "hit3d_les_combine": [_G, _P, _P, _P, _P, _D, _P, _P],
    "hit3d_rk4_combine": [_G, _I32, _D, _P, _P, _P, _P, _P, _P, _P],
}

PASS_NAMES = ("x_inverse_curl", "y_inverse_curl", "z_c2r_cross_r2c", "y_forward",
              "x_forward_stage", "fp64_finalize", "setup")

_lib = None
_lock = threading.Lock()


class Hit3dError(RuntimeError):
    """A C-ABI call returned a non-zero status."""


def load():
    """Load libhit3d_hip.so (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"{LIB_PATH} is missing: build it with `python 3dhit_amd/build.py`")
            lib = ctypes.CDLL(LIB_PATH)
            for name, args in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.argtypes = args
                fn.restype = ctypes.c_int
            _lib = lib
    return _lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(4096)
    load().hit3d_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def call(name, *args):
    rc = getattr(load(), name)(*args)
    if rc != HIT3D_OK:
        msg = f"{name}: {last_error()}"
        if rc in (HIT3D_ERR_INVALID, HIT3D_ERR_WORKSPACE):
            raise ValueError(msg)
        raise Hit3dError(msg)
    return rc
