"""The production split kernels at the full single-grid workload shapes: the 240 x 120 x 100 Si
strip waveguide (library grid 120 x 100 x 240, 8-bit material ids) and the 620 x 200 x 100 linear
taper (library grid 200 x 100 x 620, fill-fraction raster, 16-bit material ids), 20 nm cells.

* 20 steps from seeded random fields under the 10-cell CPML against the FP64 oracle on the same
  inputs (relative L2 <= 1e-4 over all cells and components): the oracle finishes these sizes
  in seconds.
* Energy of the closed (PEC-walled) box over 300 steps, a size-independent property: with
  E^s = E^{s-1} + A C H^{s-1/2} and H^{s+1/2} = H^{s-1/2} - B C^T E^s (A = dt/(eps dx),
  B = dt/(mu dx), C the backward-difference curl on the interior cells), the quantity
  W_s = sum eps E^s.E^s + sum mu H^{s-1/2}.H^{s+1/2} is conserved exactly in exact arithmetic; a
  wrong sign, index or coefficient ratio in either update breaks it.  The box is run through the
  PML kernel variant with zero-width layers (no slab cells), i.e. the production kernels.
"""
import importlib
import math
import sys

import numpy as np
import pytest

from oracle import oracle_py as O

pytestmark = pytest.mark.gpu

fd = importlib.import_module("3d_fdtd_amd")
DISC = 0.02
DT = 0.99 * DISC / math.sqrt(3)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def strip():
    I, J, K = 120, 100, 240
    ep = np.full((K, J, I, 3), 1.44 ** 2, np.float32)
    ep[:, J // 2 - 5:J // 2 + 6, I // 2 - 12:I // 2 + 13, :] = 3.48 ** 2
    return ep.reshape(-1), I, J, K


def taper():
    design = importlib.import_module("3d_fdtd_amd.design")
    fam = design.TaperFamily(disc=DISC, I=200, J=100, core_cells=11, npml=10, port=48, steps=10,
                             lam=(1.55, 1.55), nfreq=1)
    ep, K = fam.permittivity(np.concatenate([np.linspace(0.5, 3.0, 8), [10.0]]))
    return ep, fam.I, fam.J, K


CASES = {"strip": strip, "taper": taper}


def rel_l2(got, want):
    return float(np.linalg.norm(got - want) / np.linalg.norm(want))


@pytest.mark.parametrize("case", ["strip", "taper"])
def test_fullsize_cpml_steps_match_oracle(case):
    ep, I, J, K = CASES[case]()
    say("full-size CPML steps vs oracle:", case, I, J, K)
    f0 = np.random.default_rng(11).uniform(-1, 1, 6 * I * J * K).astype(np.float32).astype(np.float64)
    g = fd.DevGrid(I, J, K, DISC, ep)
    g.cpml((10,) * 6, 3, 1e-8, 1.0, 0.05)
    g.set_fields(f0)
    g.run(DT, 0, 20)
    got = g.fields()
    g.free()
    o = O.Grid(I, J, K, DISC)
    o.ep[:] = ep
    o.f[:] = f0
    cp = O.Cpml(o, (10,) * 6, DT, order=3, R0=1e-8, kappa_max=1.0, alpha_max=0.05)
    for _ in range(20):
        O.step_cpml(o, cp, DT)
    err = rel_l2(got, o.f)
    say(case, "rel L2 vs FP64 oracle after 20 steps:", err)
    assert err <= 1e-4


def energy(ep, f_prev, f_now, shape):
    K, J, I = shape
    e = f_now.reshape(K, J, I, 6)[..., :3]
    h0 = f_prev.reshape(K, J, I, 6)[..., 3:]
    h1 = f_now.reshape(K, J, I, 6)[..., 3:]
    return float(np.sum(ep.reshape(K, J, I, 3) * e * e) + np.sum(h0 * h1))


@pytest.mark.parametrize("case", ["strip", "taper"])
def test_fullsize_closed_box_conserves_energy(case):
    ep, I, J, K = CASES[case]()
    say("full-size closed-box energy:", case, I, J, K)
    f0 = np.random.default_rng(12).uniform(-1, 1, (K, J, I, 6)).astype(np.float32).astype(np.float64)
    f0[[0, -1]] = 0  # the boundary ring stays zero (never updated)
    f0[:, [0, -1]] = 0
    f0[:, :, [0, -1]] = 0
    g = fd.DevGrid(I, J, K, DISC, ep)
    g.cpml((0,) * 6, 3, 1e-8, 1.0, 0.05)  # zero-width layers: PML kernel variant, PEC-closed box
    g.set_fields(f0.reshape(-1))
    W = []
    done = 0
    for s in (1, 100, 300):
        g.run(DT, done, s - 1 - done)
        prev = g.fields()
        g.run(DT, s - 1, 1)
        now = g.fields()
        done = s
        W.append(energy(ep, prev, now, (K, J, I)))
        say(case, "step", s, "W", W[-1])
    g.free()
    drift = max(abs(w - W[0]) for w in W) / abs(W[0])
    say(case, "relative energy drift over 300 steps:", drift)
    assert drift <= 1e-5
