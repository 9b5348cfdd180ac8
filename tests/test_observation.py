"""Host-side observation layer without a GPU: the HDF5 key schema, run
labels and reprs, callback / snapshot scheduling and the progress line
(reference tests/test_report_schema_labels.py:28-117 and
tests/test_callbacks.py:64-106, whose expected values are used here)."""

import math
import os

import numpy as np
import pytest
import torch

hit3d = pytest.importorskip("hit3d")
from hit3d import (BandForcing, Callback, Diagnostic, FieldWriter, LinearForcing, MemorySink,
                   NoForcing, NSParams, label, make_grid, run_label)
from hit3d.integrators import _fmt_secs, _print_progress
from hit3d.schema import grid_key, is_step_key, series_field, series_key, step_key, step_num


def test_schema():
    assert step_key(50) == "step_00000050"
    assert step_num("step_00000050") == 50
    assert is_step_key("step_00000050")
    assert not is_step_key("grid")
    assert grid_key("Nx") == "grid/Nx"
    assert series_key("energy_budget") == "series/energy_budget"
    assert series_field("energy_budget", "E") == "series/energy_budget/E"


def test_labels_and_reprs():
    g = make_grid(16, dtype=torch.float64)
    assert label(g) == "N16"
    assert label(make_grid(16, 8, 4)) == "N16x8x4"
    p = NSParams(nu=1e-3)
    assert label(p) == "NavierStokes_nu0.001_NoForcing"
    f = BandForcing(eps=0.1, kmin=1.0, kmax=2.5)
    assert label(f) == "BandForcing_eps0.1_k1-2.5"
    assert label(NSParams(nu=1e-3, forcing=f)) == "NavierStokes_nu0.001_BandForcing_eps0.1_k1-2.5"
    lf = LinearForcing(A=0.1333)
    assert label(lf) == "LinearForcing_A0.1333"
    assert repr(lf) == "LinearForcing(A=0.1333)"
    assert run_label(g, p) == "N16_NavierStokes_nu0.001_NoForcing_RK4"
    assert label(NSParams(nu=np.float32(1e-3))) == "NavierStokes_nu0.001_NoForcing"
    assert repr(g) == "Grid[float64](16x16x16, L=(6.28319, 6.28319, 6.28319))"
    assert repr(NoForcing()) == "NoForcing()"
    assert repr(BandForcing(eps=0.1, kmin=1.0, kmax=2.5)) == "BandForcing(eps=0.1, k=[1, 2.5])"
    rp = repr(NSParams(nu=1e-3, forcing=f))
    assert rp == "NavierStokes(nu=0.001, forcing=BandForcing(eps=0.1, k=[1, 2.5]))"


def _due_steps(cb, dt, nsteps, t0=0.0):
    return [n for n in range(nsteps + 1) if cb.is_due(n, t0 + n * dt, dt)]


def test_fieldwriter_time_schedule():
    # no snapshot before warmup_time, then every every_time regardless of dt:
    # t = 1.0, 1.5, 2.0, 2.5, 3.0 -> steps 10, 15, 20, 25, 30
    w = FieldWriter(sink=MemorySink(), every_time=0.5, warmup_time=1.0)
    assert _due_steps(w, 0.1, 30) == [10, 15, 20, 25, 30]
    # float32 accumulated times must not count a crossing twice
    w = FieldWriter(sink=MemorySink(), every_time=0.1)
    dt = np.float32(0.01)
    t, due = np.float32(0.0), []
    for n in range(101):
        if w.is_due(n, float(t), float(dt)):
            due.append(n)
        t = np.float32(t + dt)
    assert due == list(range(0, 101, 10))


def test_step_schedules():
    assert _due_steps(FieldWriter(sink=MemorySink(), every=50), 1e-2, 100) == [0, 50, 100]
    assert _due_steps(Diagnostic(lambda s: 0.0, every=10), 1e-2, 100) == list(range(0, 101, 10))
    # time-based Callback: step 0 plus each crossing of a multiple of 0.25
    due = _due_steps(Callback(lambda s: None, every_time=0.25), 1e-2, 100)
    assert due[0] == 0 and 4 <= len(due) <= 6
    assert all(math.floor(n * 1e-2 / 0.25) > math.floor((n - 1) * 1e-2 / 0.25) for n in due[1:])


def test_fieldwriter_overwrite_semantics(tmp_path):
    path = str(tmp_path / "sub" / "snap.h5")
    os.makedirs(os.path.dirname(path))
    open(path, "w").close()
    FieldWriter(path, sink=MemorySink(), overwrite=False)
    assert os.path.exists(path)
    FieldWriter(path, sink=MemorySink())          # default overwrite=True deletes
    assert not os.path.exists(path)
    FieldWriter(str(tmp_path / "new" / "dir" / "s.h5"), sink=MemorySink())
    assert os.path.isdir(tmp_path / "new" / "dir")   # parent directories are created


def test_diagnostic_path_without_name_errors():
    with pytest.raises(ValueError):
        Diagnostic(lambda s: 0.0, path="somewhere.h5")


def test_hdf5_layer_fails_loudly_without_h5py(tmp_path):
    try:
        import h5py  # noqa: F401
        pytest.skip("h5py is installed here")
    except ImportError:
        pass
    with pytest.raises(ImportError, match="h5py"):
        FieldWriter(str(tmp_path / "a.h5"))
    with pytest.raises(ImportError, match="h5py"):
        Diagnostic(lambda s: 0.0, path=str(tmp_path / "a.h5"), name="E")
    with pytest.raises(ImportError, match="h5py"):
        hit3d.load_run(str(tmp_path / "a.h5"))


def test_progress_line(capsys):
    import time
    _print_progress(5, 10, 0.005, time.time() - 2.0)
    _print_progress(10, 10, 0.01, time.time() - 4.0)
    out = capsys.readouterr().out
    assert out.count("\n") == 2
    assert "step 10/10 (100%)" in out and "step  5/10 ( 50%)" in out
    assert "t = " in out and "eta " in out
    assert _fmt_secs(5.04) == "5.0s" and _fmt_secs(125.0) == "2m05s"


def test_gauss_newton_calibration_step():
    """The LES closure calibration's solver (host logic): on a residual that
    is linear in (log gain, tilt) it lands on the root after one clamped
    step; the best point is returned and every evaluation is recorded."""
    les_experiment = pytest.importorskip("hit3d.les_experiment")
    k = np.arange(5, 22, dtype=np.float64)
    truth = np.array([0.07, -0.16])
    A = np.stack([-np.ones_like(k), -(k / 21.0 - 0.5)], axis=1)   # more gain/tilt -> less energy

    def resid(x):
        return A @ (np.asarray(x) - truth)
    best, pts = les_experiment.gauss_newton(resid, [0.0, 0.0], (0.1, 0.3), (0.2, 0.6), 2)
    assert len(pts) == 7
    assert np.allclose(best, truth, atol=1e-12)
    # a step larger than the clamp is cut to it
    best2, pts2 = les_experiment.gauss_newton(resid, [0.0, 0.0], (0.1, 0.3), (0.02, 0.6), 1)
    assert abs(pts2[-1][0][0] - 0.02) < 1e-12


# --- report readers over a memory-held run (reference tests/test_report_schema_labels.py:45-81) ---


def _recorded_run(with_meta=True, dtype=np.complex64):
    sink = MemorySink()
    meta = {"Nx": 16, "Ny": 8, "Nz": 32, "Lx": 3.0, "Ly": 2.0, "Lz": 1.0, "nu": 0.01}
    rng = np.random.default_rng(0)
    for step in (50, 0, 100):                      # written out of order
        u = (rng.standard_normal((3, 16, 8, 17)) + 1j * rng.standard_normal((3, 16, 8, 17))).astype(dtype)
        sink.write(step, step * 1e-3, {"u_hat": u}, meta if with_meta else {})
    if not with_meta:
        sink.meta = None
    return sink


def test_load_run_from_memory_sink():
    from hit3d.report import read_grid, read_snapshot, step_keys
    sink = _recorded_run()
    run = hit3d.load_run(sink)
    assert run.steps == [0, 50, 100] and run.nu == 0.01 and run.dtype == torch.float32
    g = run.grid
    assert (g.Nx, g.Ny, g.Nz, g.Lx, g.Ly, g.Lz) == (16, 8, 32, 3.0, 2.0, 1.0)
    assert step_keys(sink) == ["step_00000000", "step_00000050", "step_00000100"]
    assert np.array_equal(read_snapshot(sink, 50), sink.snapshots[50]["u_hat"])
    g64, nu = read_grid(sink, torch.float64)
    assert g64.dtype == torch.float64 and nu == 0.01
    assert hit3d.load_run(_recorded_run(dtype=np.complex128)).dtype == torch.float64
    assert hit3d.read_series(sink, "energy_budget") is None


def test_load_run_without_grid_metadata_warns_and_assumes_2pi_cube():
    sink = _recorded_run(with_meta=False)
    with pytest.warns(UserWarning, match="no grid metadata"):
        run = hit3d.load_run(sink)
    g = run.grid
    assert (g.Nx, g.Ny, g.Nz) == (16, 8, 32) and run.nu is None
    assert abs(g.Lx - 2 * math.pi) < 1e-15


def test_load_run_of_an_empty_run_errors():
    with pytest.raises(ValueError, match="no snapshots found"):
        hit3d.load_run(MemorySink())


def test_series_round_trip_through_a_memory_sink():
    sink = MemorySink()
    d = Diagnostic(lambda s: 0.0, every=1, path=sink, name="E")
    d.times, d.values = [0.0, 0.1, 0.2], [1.0, 0.9, 0.8]
    d.finalize()
    b = Diagnostic(lambda s: 0.0, every=1, path=sink, name="energy_budget")
    b.times, b.values = [0.0, 0.5], [{"E": 1.0, "eps": 0.1, "P": 0.0}, {"E": 0.9, "eps": 0.2, "P": 0.0}]
    b.finalize()
    s = hit3d.read_series(sink, "E")
    assert set(s) == {"t", "values"} and np.allclose(s["values"], [1.0, 0.9, 0.8])
    s = hit3d.read_series(sink, "energy_budget")
    assert set(s) == {"t", "E", "eps", "P"} and np.allclose(s["eps"], [0.1, 0.2])
    with pytest.raises(ValueError, match="already exists"):
        hit3d.save_series(sink, "E", d)
