"""evolve with callbacks on the GPU (reference tests/test_callbacks.py:30-97,
tests/test_report_schema_labels.py:43-88): diagnostics, time-scheduled
callbacks, asynchronous snapshots (checked against independent rollouts),
progress output, and the HDF5 files where h5py is installed."""

import numpy as np
import pytest
import torch

import hit3d_oracle as O

pytestmark = pytest.mark.gpu

h3 = pytest.importorskip("hit3d")

N, NU = 32, 0.05


def beltrami(g):
    og = O.make_grid(g.Nx, g.Ny, g.Nz, dtype=np.float64)
    return torch.from_numpy(O.beltrami(og)).to(g.complex_dtype).cuda()


def run_beltrami(g, callbacks, dt=1e-2, nsteps=100, progress=False):
    return h3.evolve(h3.make_ns_rhs(g), beltrami(g), h3.NSParams(nu=NU), g, dt, nsteps,
                     callbacks=callbacks, progress=progress)


def test_diagnostic_callback_and_snapshot_scheduling():
    g = h3.make_grid(N, dtype=torch.float64)
    dt, nsteps = 1e-2, 100
    E = h3.Diagnostic(lambda s: h3.energy(s.u_hat, s.grid), every=10)
    ncalls = []
    probe = h3.Callback(lambda s: ncalls.append(s.step), every_time=0.25)
    sink = h3.MemorySink()
    writer = h3.FieldWriter(sink=sink, every=50)
    u = run_beltrami(g, (E, probe, writer), dt, nsteps)
    # Diagnostic: step 0 and every 10th step, energy decays as exp(-2 nu t)
    assert len(E.times) == nsteps // 10 + 1
    assert np.isclose(E.times[1], 10 * dt)
    expected = E.values[0] * np.exp(-2 * NU * np.asarray(E.times))
    assert np.allclose(np.asarray(E.values), expected, rtol=1e-8)
    # time-based callback: the initial state plus crossings of 0.25
    assert 4 <= len(ncalls) <= 6
    # snapshots at 0, 50, 100; the last is the evolved state; grid metadata
    assert sorted(sink.snapshots) == [0, 50, 100]
    assert np.array_equal(sink.snapshots[100]["u_hat"], u.cpu().numpy())
    assert np.isclose(sink.snapshots[100]["t"], dt * nsteps)
    assert sink.meta["Nx"] == N and np.isclose(sink.meta["Lx"], 2 * np.pi)
    assert sink.meta["nu"] == NU


def test_async_snapshots_match_independent_rollouts():
    """Every step of a 64^3 float32 run is copied out while the solver goes
    on (pinned buffers reused, at most two in flight); each copy equals a
    separate rollout to that step, bit for bit."""
    g = h3.make_grid(64, dtype=torch.float32)
    rhs = h3.make_ns_rhs(g)
    params = h3.NSParams(nu=1 / 400)
    u0 = h3.random_field(3, g, h3.config_spectrum)
    sink = h3.MemorySink()
    uN = h3.evolve(rhs, u0, params, g, 2e-3, 12,
                   callbacks=(h3.FieldWriter(sink=sink, every=1, depth=2),))
    assert sorted(sink.snapshots) == list(range(13))
    for n in (0, 1, 5, 12):
        ref = h3.rollout(rhs, u0, params, 2e-3, n).cpu().numpy()
        assert np.array_equal(sink.snapshots[n]["u_hat"], ref), n
        assert np.isclose(sink.snapshots[n]["t"], 2e-3 * n)
    assert np.array_equal(sink.snapshots[12]["u_hat"], uN.cpu().numpy())


def test_sink_errors_surface_at_finalize():
    class Broken:
        def write(self, step, t, arrays, meta):
            raise OSError("disk full")

    g = h3.make_grid(16, dtype=torch.float64)
    with pytest.raises(OSError, match="disk full"):
        run_beltrami(g, (h3.FieldWriter(sink=Broken(), every=5),), nsteps=10)


def test_progress_printing(capsys):
    g = h3.make_grid(N, dtype=torch.float64)
    run_beltrami(g, (), dt=1e-3, nsteps=10, progress=5)
    out = capsys.readouterr().out
    assert out.count("\n") == 2
    assert "step 10/10 (100%)" in out and "t = " in out and "eta " in out
    run_beltrami(g, (), dt=1e-3, nsteps=5)
    assert capsys.readouterr().out == ""
    run_beltrami(g, (), dt=1e-3, nsteps=5, progress=True)
    out = capsys.readouterr().out
    assert out and "(100%)" in out


def test_hdf5_files(tmp_path):
    h5py = pytest.importorskip("h5py")
    g = h3.make_grid(N, dtype=torch.float64)
    path = str(tmp_path / "snap.h5")
    E = h3.Diagnostic(lambda s: h3.energy(s.u_hat, s.grid), every=10, path=path, name="E")
    u = run_beltrami(g, (h3.FieldWriter(path, every=50), E))
    with h5py.File(path, "r") as f:
        steps = sorted(k for k in f if k.startswith("step_"))
        assert steps == ["step_00000000", "step_00000050", "step_00000100"]
        assert np.allclose(f["step_00000100/u_hat"][()], u.cpu().numpy())
        assert f["grid/Nx"][()] == N and f["grid/nu"][()] == NU
    run = h3.load_run(path)
    assert run.steps == [0, 50, 100] and run.nu == NU and run.dtype == torch.float64
    s = h3.read_series(path, "E")
    assert np.allclose(s["t"], E.times) and np.allclose(s["values"], E.values)
    assert h3.read_series(path, "energy_budget") is None


def test_recorded_run_read_back_through_the_report_readers():
    """evolve with FieldWriter and Diagnostic both recording into one
    MemorySink; load_run / read_snapshot / read_series read the run back
    (the same readers serve HDF5 files)."""
    from hit3d.report import read_snapshot
    g = h3.make_grid(N, dtype=torch.float64)
    sink = h3.MemorySink()
    E = h3.Diagnostic(lambda s: h3.energy(s.u_hat, s.grid), every=10, path=sink, name="E")
    B = h3.Diagnostic(h3.energy_budget_cb, every=25, path=sink, name="energy_budget")
    u = run_beltrami(g, (h3.FieldWriter(sink=sink, every=50), E, B))
    run = h3.load_run(sink)
    assert run.steps == [0, 50, 100] and run.nu == NU and run.dtype == torch.float64
    assert (run.grid.Nx, run.grid.Ny, run.grid.Nz) == (N, N, N)
    assert np.array_equal(read_snapshot(sink, 100), u.cpu().numpy())
    s = h3.read_series(sink, "E")
    assert np.allclose(s["t"], E.times) and np.allclose(s["values"], E.values)
    b = h3.read_series(sink, "energy_budget")
    assert set(b) == {"t", "E", "eps", "P"} and len(b["t"]) == 5
    assert np.allclose(b["eps"], 2 * NU * np.asarray(b["E"]), rtol=1e-10)   # Beltrami: Omega = E
