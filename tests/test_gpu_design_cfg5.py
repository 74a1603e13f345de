"""GPU test of the surrogate design loop on the batched taper set's family at reduced counts: 20 nm cells,
200 x 100 cross-section, 8-cell CPML, 6000 steps, 16 wavelengths, TE0 port modes from the mode
solver (mode_solve_beta_squared), the fill-fraction raster on the device.  Reduced: lengths
U[3, 4] um, fixed end widths (1.0 and 2.0 um, so two port modes), 32 training designs, 2000
candidates screened and 8 verified per iteration, two iterations.

Checks: the loop accepts only designs whose FDTD band |t|^2 reaches the spec, knows the surrogate's
error on each, and the GPU FDTD transmission of two verified designs equals the FP64 oracle's to
|dT| <= 1e-3 at every wavelength (measured about 1e-7)."""
import importlib
import sys

import numpy as np
import pytest

from oracle import taper_oracle

pytestmark = pytest.mark.gpu

fd = importlib.import_module("3d_fdtd_amd")
design = importlib.import_module("3d_fdtd_amd.design")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_cfg5_loop_reduced_counts_and_oracle():
    fam = design.TaperFamily(disc=0.02, I=200, J=100, core_cells=11, npml=8, port=0, wmin=0.4, wmax=3.2, lmin=3.0,
                             lmax=4.0, steps=6000, lam=(1.50, 1.60), nfreq=16, ports="te0", fixed_ports=(1.0, 2.0))
    rng = np.random.default_rng(55)
    train = fam.random(rng, 32)
    say("cfg5 reduced: solving the two port modes, simulating 32 training designs")
    fam.precompute_ports(train, workers=2)
    t_train = fam.simulate(train)
    T_train = design.band_transmission(t_train)
    spec = float(np.quantile(T_train, 0.6))
    say("training band T: min %.3f median %.3f max %.3f; spec %.3f" % (T_train.min(), np.median(T_train),
                                                                        T_train.max(), spec))
    net = design.fit(fam, train, t_train, steps=300)
    loop = design.SurrogateLoop(fam, net, train, t_train, spec=spec, pool=2000, verify=8, refit=5, min_hit=0.0,
                                seed=56)
    loop.run(iterations=2)
    say("accepted %d of %d verified" % (len(loop.accepted), loop.verified))
    assert loop.verified == 16
    assert all(T >= spec for T in loop.T_fdtd)
    err = loop.surrogate_error()
    assert len(err) == len(loop.accepted)
    net.free()
    # GPU FDTD vs the FP64 oracle on two designs of the training set
    for q in (int(np.argmax(T_train)), int(np.argmin(T_train))):
        say("oracle check of training design", q)
        t_cpu = taper_oracle.transmission(**fam.run_inputs(train[q]))
        dT = np.abs(np.abs(t_cpu) ** 2 - np.abs(t_train[q]) ** 2).max()
        dt_ = np.abs(t_cpu - t_train[q]).max()
        say("  max |dT| %.2e, max |dt| %.2e, band T %.4f" % (dT, dt_, float(np.mean(np.abs(t_cpu) ** 2))))
        assert dT <= 1e-3 and dt_ <= 1e-3
