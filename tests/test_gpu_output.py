"""GPU side of the drop-in output formats: the field and Fourier slice dumps of
timefield3d_plot_*_slice (timefield3d.c:348-1188 data layout) read the device state, so they
must equal the fields and running DFT that the same grid hands back through the drop-in API
(FP32 fields widened to FP64, printed with 6 decimals)."""
import importlib
import sys

import numpy as np
import pytest

from tests.test_output_formats import _read_matrix

pytestmark = pytest.mark.gpu

fd = importlib.import_module("3d_fdtd_amd")
I, J, K, DISC = 26, 18, 22, 0.02


def test_field_and_fourier_plots_match_device_state(tmp_path, monkeypatch):
    print("field/fourier slice dumps", I, J, K, flush=True)
    sys.stdout.flush()
    monkeypatch.setenv("FDTD3D_PLOT_DIR", str(tmp_path))
    L = fd.lib()
    g = fd.Grid(I, J, K, DISC)
    g.a["ep"][:] = 2.0
    g.a["mu"][:] = 1.0
    g.a["fields"][:] = np.random.default_rng(5).uniform(-1, 1, g.a["fields"].size)
    g.push()
    dt = 0.99 * DISC / np.sqrt(3)
    for n in range(4):
        g.step(dt)
        L.timefield3d_update_fourier_fields(g.tf, 0.3 * n)
    L.timefield3d_plot_Ex_y_slice(g.tf, 7)
    L.timefield3d_plot_Ey_z_slice(g.tf, 11)
    L.timefield3d_plot_Ez_x_slice(g.tf, 13)
    L.timefield3d_plot_Ex_z_slice_fourier(g.tf, 9)
    a = g.pull()
    f = a["fields"].reshape(K, J, I, 6)
    four = a["fourier"].reshape(K, J, I, 6, 2)
    cases = [("Ex_y_7", f[:, 7, :, 0].T), ("Ey_z_11", f[11, :, :, 1]), ("Ez_x_13", f[:, :, 13, 2].T),
             ("Ex_fourier_z_9", np.hypot(four[9, :, :, 0, 0], four[9, :, :, 0, 1]))]
    for name, want in cases:
        ncol, cols, rows, vals = _read_matrix(tmp_path / f"{name}.txt")
        assert vals.shape == want.shape and ncol == want.shape[1], name
        err = np.abs(vals - want).max()
        print(name, "max |dump - state|", err, flush=True)
        assert err <= 5e-7 + 1e-6 * np.abs(want).max()
        assert np.abs(want).max() > 0
    g.free()
