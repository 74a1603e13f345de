"""GPU tests of the DFT-domain permittivity gradient: box DFT monitors, the overlap monitor's
adjoint source and the gradient kernel, through the C-ABI.

* Parity: forward and adjoint device runs vs the FP64 oracle on the same inputs (monitor value,
  both DFT boxes, the gradient) to relative 1e-4.
* Physics: the device gradient equals central differences of device forward runs for eps changes
  on three 2x2x2 blocks (relative 1.5%; the oracle's own check, with a planted mistake, is in
  test_oracle_invariants.py).
"""
import ctypes as C
import importlib
import sys

import numpy as np
import pytest

from tests import gradient_case as GC
from tests.test_oracle_invariants import oracle_gradient

pytestmark = pytest.mark.gpu

fd = importlib.import_module("3d_fdtd_amd")
L = fd.lib
DP = C.POINTER(C.c_double)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _grid(ep):
    g = fd.Grid(GC.I, GC.J, GC.K, GC.DISC)
    g.a["ep"][:] = ep
    g.push()
    npml = (C.c_int * 6)(*GC.NPML)
    fd.check(L().fdtd3d_ext_set_cpml(g.tf, npml, GC.CPML["order"], GC.CPML["R0"], GC.CPML["kappa_max"],
                                     GC.CPML["alpha_max"]), "cpml")
    return g


def _schedule():
    n = np.arange(GC.NSTEPS)
    return GC.envelope(n).astype(np.float64), (GC.OMEGA * n * GC.DT).astype(np.float64)


def _dft_box(g):
    box = (C.c_int * 6)(*GC.BOX)
    om = np.array([GC.OMEGA])
    bid = L().fdtd3d_ext_run_add_dft_box(g.tf, box, 7, 1, om.ctypes.data_as(DP))
    assert bid >= 0
    return bid


def device_forward(ep, with_box):
    g = _grid(ep)
    md = fd.mode_from_array(GC.MI, GC.MJ, GC.DISC, GC.mode_array())
    amp, phase = _schedule()
    assert L().fdtd3d_ext_run_add_mode_source(g.tf, GC.I // 2, GC.J // 2, GC.SRC_K, md, 0, 0, GC.NSTEPS,
                                              amp.ctypes.data_as(DP), phase.ctypes.data_as(DP)) >= 0
    om = np.array([GC.OMEGA])
    mid = L().fdtd3d_ext_run_add_overlap(g.tf, GC.I // 2, GC.J // 2, GC.MON_K, md, 0, 1, om.ctypes.data_as(DP))
    assert mid >= 0
    bid = _dft_box(g) if with_box else None
    fd.check(L().fdtd3d_ext_run(g.tf, GC.DT, 0, GC.NSTEPS, 0.0), "forward run")
    cum = fd.Complex()
    fd.check(L().fdtd3d_ext_run_overlap(g.tf, mid, C.byref(cum), None), "overlap")
    return g, complex(cum.re, cum.im), bid


def device_adjoint(ep):
    g = _grid(ep)
    md = fd.mode_from_array(GC.MI, GC.MJ, GC.DISC, GC.mode_array())
    amp, phase = _schedule()
    assert L().fdtd3d_ext_run_add_adjoint_source(g.tf, GC.I // 2, GC.J // 2, GC.MON_K, md, 0, GC.OMEGA * GC.DT,
                                                 GC.NSTEPS, amp.ctypes.data_as(DP), phase.ctypes.data_as(DP)) >= 0
    bid = _dft_box(g)
    fd.check(L().fdtd3d_ext_run(g.tf, GC.DT, 0, GC.NSTEPS, 1.0), "adjoint run")
    return g, bid, float(amp.sum())


def box_values(g, bid):
    nb = GC.box_cells()
    out = (fd.Complex * (3 * nb))()
    fd.check(L().fdtd3d_ext_run_dft(g.tf, bid, out), "dft box")
    return np.ctypeslib.as_array(C.cast(out, DP), shape=(3 * nb * 2,)).copy()


def device_gradient(gf, bf, ga, ba, E0):
    w = GC.gradient_weight(E0)
    nb = GC.box_cells()
    gr, gi = np.zeros(nb), np.zeros(nb)
    fd.check(L().fdtd3d_ext_dft_gradient(gf.tf, bf, ga.tf, ba, 0, fd.Complex(w.real, w.imag),
                                         gr.ctypes.data_as(DP)), "gradient")
    fd.check(L().fdtd3d_ext_dft_gradient(gf.tf, bf, ga.tf, ba, 0, fd.Complex(w.imag, -w.real),
                                         gi.ctypes.data_as(DP)), "gradient")
    i0, i1, j0, j1, k0, k1 = GC.BOX
    return (gr + 1j * gi).reshape(k1 - k0, j1 - j0, i1 - i0)


def rel(got, want):
    return float(np.linalg.norm(np.asarray(got) - np.asarray(want)) / np.linalg.norm(want))


@pytest.fixture(scope="module")
def device_case():
    ep = GC.materials()
    say("gradient case: forward + adjoint on the device")
    gf, a, bf = device_forward(ep, True)
    ga, ba, E0 = device_adjoint(ep)
    grad = device_gradient(gf, bf, ga, ba, E0)
    F, A = box_values(gf, bf), box_values(ga, ba)
    gf.free()
    ga.free()
    return a, grad, F, A


def test_gradient_parity_with_oracle(device_case):
    a, grad, F, A = device_case
    say("gradient case: forward + adjoint in the oracle")
    a_o, grad_o, F_o, A_o, _ = oracle_gradient()
    errs = (abs(a - a_o) / abs(a_o), rel(F, F_o), rel(A, A_o), rel(grad, grad_o))
    say("monitor %.2e forward box %.2e adjoint box %.2e gradient %.2e" % errs)
    assert max(errs) <= 1e-4


def test_gradient_matches_device_finite_differences(device_case):
    _, grad, _, _ = device_case
    ep = GC.materials()
    i0, i1, j0, j1, k0, k1 = GC.BOX
    delta = 0.2  # larger than the oracle's 0.05: FP32 rounding of the monitor must stay small against the change
    for corner in GC.BLOCKS:
        mask = GC.block_mask(corner)
        vals = []
        for sgn in (1, -1):
            e = ep.reshape(GC.K, GC.J, GC.I, 3).copy()
            e[k0:k1, j0:j1, i0:i1][mask] += np.float32(sgn * delta)
            g, a, _ = device_forward(e.reshape(-1), False)
            g.free()
            vals.append(a)
        fdv = (vals[0] - vals[1]) / (2 * delta)
        err = abs(grad[mask].sum() - fdv) / abs(fdv)
        say("block", corner, "device fd", fdv, "device adjoint", grad[mask].sum(), "rel err", err)
        assert err < 0.015
