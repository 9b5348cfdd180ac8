"""hit3d_energy_dissipation_noalloc (the allocation-free form, scratch in a
device array of the library) against hit3d_energy_dissipation (caller
workspace) through the C ABI: bit-identical FP64 results on two calls in a
row, the second on a smaller grid, in float32 and float64."""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

h3 = pytest.importorskip("hit3d")


def _both(u, g, nu, kmin, kmax):
    from hit3d import _lib
    from hit3d.grid import stream
    ws = g.workspace(_lib.OP_REDUCE)
    a = torch.full((4,), float("nan"), dtype=torch.float64, device=u.device)
    b = torch.full((4,), float("nan"), dtype=torch.float64, device=u.device)
    _lib.call("hit3d_energy_dissipation", g.desc, u.data_ptr(), nu, kmin, kmax, a.data_ptr(),
              ws.data_ptr(), ws.numel(), stream())
    _lib.call("hit3d_energy_dissipation_noalloc", g.desc, u.data_ptr(), nu, kmin, kmax, b.data_ptr(),
              stream())
    torch.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_noalloc_energy_dissipation_matches_workspace_form(dtype):
    for n, seed, nu, kmin, kmax in ((128, 1, 1e-3, 2.0, 5.0), (32, 2, 0.05, 1.0, 3.0)):
        g = h3.make_grid(n, dtype=dtype)
        u = h3.random_field(seed, g, h3.config_spectrum)
        a, b = _both(u, g, nu, kmin, kmax)
        assert np.all(np.isfinite(a)) and a[0] > 0 and a[3] > 0
        assert np.array_equal(a, b), (n, a, b)
        assert abs(b[0] - h3.energy(u, g)) == 0.0


def test_noalloc_rejects_bad_arguments():
    from hit3d import _lib
    g = h3.make_grid(16)
    out = torch.empty(4, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        _lib.call("hit3d_energy_dissipation_noalloc", g.desc, None, 0.0, 1.0, 2.0, out.data_ptr(),
                  None)
    bad = _lib.GridDesc(12, 16, 16, 6.28, 6.28, 6.28, 4, 0)
    u = torch.zeros((3, 16, 16, 9), dtype=torch.complex64, device="cuda")
    with pytest.raises(ValueError):
        _lib.call("hit3d_energy_dissipation_noalloc", ctypes.byref(bad), u.data_ptr(), 0.0, 1.0, 2.0,
                  out.data_ptr(), None)
