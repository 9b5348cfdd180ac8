"""CPU-side checks of the C ABI: the library loads, exports every function
include/hit3d.h declares, and rejects bad arguments before touching the GPU."""

import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "hit3d.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^int\s+(hit3d_\w+)\s*\(", text, re.M)))


def _lib():
    from hit3d import _lib
    try:
        return _lib.load()
    except ImportError as e:
        pytest.fail(f"library not built: {e}")


def test_header_declares_entry_points():
    names = _declared()
    assert "hit3d_rk4_steps" in names and "hit3d_shell_spectrum" in names and len(names) >= 15


def test_library_exports_every_declared_symbol():
    lib = _lib()
    from hit3d import _lib as L
    for name in _declared():
        assert hasattr(lib, name), name
        assert name in L.SIGNATURES, f"{name} has no ctypes signature"


def test_struct_layouts_match_header():
    from hit3d import _lib as L
    assert ctypes.sizeof(L.GridDesc) == 3 * 8 + 3 * 8 + 4 + 4
    assert ctypes.sizeof(L.ForcingDesc) == 4 + 4 + 7 * 8


def _grid(nx=16, ny=16, nz=16, prec=4):
    from hit3d import _lib as L
    return L.GridDesc(nx, ny, nz, 6.283185307179586, 6.283185307179586, 6.283185307179586, prec, 0)


def test_invalid_grids_rejected_without_gpu():
    from hit3d import _lib as L
    lib = _lib()
    n = ctypes.c_size_t()
    for bad in (_grid(12), _grid(16, 2048), _grid(16, 16, 2), _grid(prec=2)):
        rc = lib.hit3d_workspace_size(ctypes.byref(bad), L.OP_RK4, 3, 1, ctypes.byref(n))
        assert rc == L.HIT3D_ERR_INVALID
        assert "grid" in L.last_error() or "precision" in L.last_error()
    assert lib.hit3d_workspace_size(ctypes.byref(_grid()), 99, 3, 1, ctypes.byref(n)) == L.HIT3D_ERR_INVALID


def test_workspace_sizes_and_null_checks():
    from hit3d import _lib as L
    lib = _lib()
    g = _grid(64, 64, 64)
    n_pruned, n_dense = ctypes.c_size_t(), ctypes.c_size_t()
    assert lib.hit3d_workspace_size(ctypes.byref(g), L.OP_RK4, 3, 1, ctypes.byref(n_pruned)) == 0
    assert lib.hit3d_workspace_size(ctypes.byref(g), L.OP_RK4, 3, 0, ctypes.byref(n_dense)) == 0
    assert 0 < n_pruned.value < n_dense.value
    nb = ctypes.c_int64()
    assert lib.hit3d_shell_bins(ctypes.byref(g), ctypes.byref(nb)) == 0 and nb.value == 57 + 1
    f = L.ForcingDesc(L.FORCING_NONE, 0)
    rc = lib.hit3d_rk4_steps(ctypes.byref(g), None, 0.1, 1e-3, 1, 0.0, ctypes.byref(f), 1, None,
                             None, 0, None)
    assert rc == L.HIT3D_ERR_INVALID and "u_hat" in L.last_error()
    buf = ctypes.create_string_buffer(8)
    full = lib.hit3d_last_error(buf, 8)
    assert full > 7 and len(buf.value) == 7
    bad = L.ForcingDesc(7, 0)
    rc = lib.hit3d_rhs(ctypes.byref(g), ctypes.c_void_p(256), ctypes.c_void_p(512), 0.1,
                       ctypes.byref(bad), 0.0, 1, None, 0, None)
    assert rc == L.HIT3D_ERR_INVALID and "forcing" in L.last_error()


def test_adjoint_record_and_forcing_term_argument_checks():
    """Argument checks of hit3d_rk4_steps_record, hit3d_rk4_step_vjp's
    stages and hit3d_forcing_term: they fail before any device work (the
    pointers below are never dereferenced)."""
    from hit3d import _lib as L
    lib = _lib()
    g = _grid(16, 16, 16)
    f = L.ForcingDesc(L.FORCING_NONE, 0)
    P = ctypes.c_void_p
    rc = lib.hit3d_rk4_steps_record(ctypes.byref(g), P(4096), 0.1, 1e-3, 2, 0.0, ctypes.byref(f), 1, None,
                                    None, 0, None)
    assert rc == L.HIT3D_ERR_INVALID and "record" in L.last_error()
    rc = lib.hit3d_rk4_steps_record(ctypes.byref(g), P(4096), 0.1, 1e-3, 2, 0.0, ctypes.byref(f), 1, P(4096),
                                    None, 0, None)
    assert rc == L.HIT3D_ERR_INVALID and "alias" in L.last_error()
    state = 3 * 16 * 16 * 9 * 8                       # one complex64 state of a 16^3 grid
    rc = lib.hit3d_rk4_step_vjp(ctypes.byref(g), P(1 << 20), P(2 << 20), P((3 << 20) + state), 0.1, 1e-3,
                                0.0, ctypes.byref(f), 1, P(3 << 20), None, None, 0, None)
    assert rc == L.HIT3D_ERR_INVALID and "stages" in L.last_error()   # ubar inside the stage inputs
    rc = lib.hit3d_forcing_term(ctypes.byref(g), P(4096), ctypes.byref(f), P(4096), None, 0, None)
    assert rc == L.HIT3D_ERR_INVALID and "alias" in L.last_error()


def test_rk4_combine_argument_checks():
    """hit3d_rk4_combine rejects a stage outside 1..4, a non-finite dt, a
    missing u for stages 1-3, acc aliasing out and a hold coefficient without
    its stage state, before any device work."""
    from hit3d import _lib as L
    lib = _lib()
    g = _grid(16, 16, 16)
    P = ctypes.c_void_p
    for stage, dt, u, hc, acc, out, word in ((0, 0.1, P(4096), None, P(8192), P(12288), "stage"),
                                             (5, 0.1, P(4096), None, P(8192), P(12288), "stage"),
                                             (1, float("nan"), P(4096), None, P(8192), P(12288), "dt"),
                                             (2, 0.1, None, None, P(8192), P(12288), "u_hat"),
                                             (3, 0.1, P(4096), None, P(8192), P(8192), "different"),
                                             (3, 0.1, P(4096), P(20480), P(8192), P(12288), "both")):
        rc = lib.hit3d_rk4_combine(ctypes.byref(g), stage, dt, u, P(16384), hc, None, acc, out, None)
        assert rc == L.HIT3D_ERR_INVALID and word in L.last_error(), (stage, L.last_error())
