"""GPU test of the DFT-domain permittivity gradient on the full-size taper grid: the linear Si
taper 0.5 -> 3.0 um over 10 um between strip ports, 620 x 200 x 100 cells of 20 nm (library axes
200 x 100 x 620, 12.4 M cells, so the field arrays exceed the last-level cache and the launches
take their streaming-load path), 10-cell CPML, TE0 port modes from the mode solver.

Forward run (mode source at the input port, overlap monitor downstream, FP64 DFT box over the core
rows of the first 110 planes of the taper region) and adjoint run (the monitor's adjoint source,
the same box), then the gradient kernel -- against the FP64 oracle on the same inputs: monitor
value, both DFT boxes and the gradient to relative 1e-4.  The runs are 1500 steps (the oracle takes
about 20 ms per step on 16 host threads); the monitor sits 150 planes from the source so that the
pulse reaches it within them.
"""
import ctypes as C
import importlib
import math
import sys

import numpy as np
import pytest

from oracle import oracle_py as O

pytestmark = pytest.mark.gpu

fd = importlib.import_module("3d_fdtd_amd")
WL = importlib.import_module("3d_fdtd_amd.workloads")
design = importlib.import_module("3d_fdtd_amd.design")

NSTEPS = 1500
NPML = 10
W0 = 2 * math.pi / 1.55
CPML = dict(order=3, R0=1e-8, kappa_max=1.0, alpha_max=0.05)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(got, want):
    return float(np.linalg.norm(np.asarray(got) - np.asarray(want)) / np.linalg.norm(want))


@pytest.fixture(scope="module")
def case():
    fam = design.TaperFamily(disc=0.02, I=200, J=100, core_cells=11, npml=NPML, port=48, steps=NSTEPS,
                             lam=(1.55, 1.55), nfreq=1)
    p = np.concatenate([np.linspace(0.5, 3.0, 8), [10.0]])
    ep, K = fam.permittivity(p)
    I, J = fam.I, fam.J
    ks, km = NPML + 1, NPML + 151
    say(f"full-size taper {I}x{J}x{K}: solving the port modes")
    m_in, mi_in, mj_in, _ = WL.port_mode(fam, p, ks, W0)
    m_out, mi_out, mj_out, _ = WL.port_mode(fam, p, km, W0)
    j0 = J // 2 - fam.core // 2
    box = (0, I, j0, j0 + fam.core, NPML + 30, NPML + 140)
    n = np.arange(NSTEPS)
    amp = np.exp(-((n - 450.0) / 120.0) ** 2 / 2)
    dt = fam.dt
    phase = W0 * n * dt
    return dict(fam=fam, ep=np.ascontiguousarray(ep, np.float32), I=I, J=J, K=K, ks=ks, km=km, m_in=m_in,
                mi_in=mi_in, mj_in=mj_in, m_out=m_out, mi_out=mi_out, mj_out=mj_out, box=box, amp=amp, phase=phase,
                dt=dt)


def nbox(c):
    i0, i1, j0, j1, k0, k1 = c["box"]
    return (i1 - i0) * (j1 - j0) * (k1 - k0)


def device_runs(c):
    I, J, K, dt = c["I"], c["J"], c["K"], c["dt"]
    om = np.array([W0])
    g = fd.DevGrid(I, J, K, 0.02, c["ep"])
    g.cpml((NPML,) * 6, CPML["order"], CPML["R0"], CPML["kappa_max"], CPML["alpha_max"])
    g.add_source(I // 2, J // 2, c["ks"], c["m_in"], c["mi_in"], c["mj_in"], 0.02, c["amp"], c["phase"])
    mon = g.add_overlap(I // 2, J // 2, c["km"], c["m_out"], c["mi_out"], c["mj_out"], 0.02, om)
    bf = g.add_dft_box(c["box"], 7, om)
    g.run(dt, 0, NSTEPS)
    a = g.overlap(mon, 1)[0][0]
    ga = fd.DevGrid(I, J, K, 0.02, c["ep"])
    ga.cpml((NPML,) * 6, CPML["order"], CPML["R0"], CPML["kappa_max"], CPML["alpha_max"])
    ga.add_adjoint_source(I // 2, J // 2, c["km"], c["m_out"], c["mi_out"], c["mj_out"], 0.02, W0 * dt, c["amp"],
                          c["phase"])
    ba = ga.add_dft_box(c["box"], 7, om)
    ga.run(dt, 0, NSTEPS, 1.0)
    nb = nbox(c)
    w = 2.0 * (np.exp(-1j * W0 * dt) - 1.0) / (dt * c["amp"].sum())
    grad = g.gradient(bf, ga, ba, 0, w, nb) + 1j * g.gradient(bf, ga, ba, 0, -1j * w, nb)
    F, A = g.dft_box(bf, nb), ga.dft_box(ba, nb)
    g.free()
    ga.free()
    return a, F, A, grad


def oracle_runs(c):
    I, J, K, dt = c["I"], c["J"], c["K"], c["dt"]
    nb = nbox(c)
    box = np.array(c["box"], np.int32)
    om = np.array([W0])
    L = O.lib()
    g = O.Grid(I, J, K, 0.02)
    g.ep[:] = c["ep"]
    cp = O.Cpml(g, (NPML,) * 6, dt, **CPML)
    acc_f = np.zeros(2 * 3 * nb)
    cum, inst = np.zeros(2), np.zeros(2)
    a = 0j
    m_in = np.ascontiguousarray(c["m_in"])
    m_out = np.ascontiguousarray(c["m_out"])
    for n in range(NSTEPS):
        L.orc_mode_source(g.ptr, I // 2, J // 2, c["ks"], c["mi_in"], c["mj_in"], 0.02, O.dptr(m_in),
                          float(c["phase"][n]), float(c["amp"][n]), 0)
        O.step_cpml(g, cp, dt)
        L.orc_mode_overlap(g.ptr, I // 2, J // 2, c["km"], c["mi_out"], c["mj_out"], 0.02, O.dptr(m_out), 0.0, 0,
                           O.dptr(cum), O.dptr(inst))
        a += np.exp(1j * W0 * n * dt) * complex(inst[0], inst[1])
        L.orc_dft_box(g.ptr, O.iptr(box), 7, 1, O.dptr(om), n * dt, O.dptr(acc_f))
    del g, cp
    g = O.Grid(I, J, K, 0.02)
    g.ep[:] = c["ep"]
    cp = O.Cpml(g, (NPML,) * 6, dt, **CPML)
    acc_a = np.zeros(2 * 3 * nb)
    for n in range(NSTEPS):
        L.orc_overlap_adjoint_source(g.ptr, I // 2, J // 2, c["km"], c["mi_out"], c["mj_out"], 0.02, O.dptr(m_out), 0,
                                     W0 * dt, float(c["amp"][n]), float(c["phase"][n]))
        O.step_cpml(g, cp, dt)
        L.orc_dft_box(g.ptr, O.iptr(box), 7, 1, O.dptr(om), (n + 1) * dt, O.dptr(acc_a))
    w = 2.0 * (np.exp(-1j * W0 * dt) - 1.0) / (dt * c["amp"].sum())
    gr, gi = np.zeros(nb), np.zeros(nb)
    L.orc_dft_gradient(nb, 7, 1, O.dptr(acc_f), 7, 1, O.dptr(acc_a), 0, w.real, w.imag, O.dptr(gr))
    L.orc_dft_gradient(nb, 7, 1, O.dptr(acc_f), 7, 1, O.dptr(acc_a), 0, w.imag, -w.real, O.dptr(gi))
    return a, acc_f, acc_a, gr + 1j * gi


def test_fullsize_taper_gradient_vs_oracle(case):
    say("full-size taper: forward + adjoint on the device")
    a, F, A, grad = device_runs(case)
    say("full-size taper: forward + adjoint in the oracle")
    a_o, F_o, A_o, grad_o = oracle_runs(case)
    errs = (abs(a - a_o) / abs(a_o), rel(F, F_o), rel(A, A_o), rel(grad, grad_o))
    say("|a| %.3e; rel errors: monitor %.2e forward box %.2e adjoint box %.2e gradient %.2e" % ((abs(a_o),) + errs))
    assert abs(a_o) > 1e-6  # the pulse reached the monitor
    assert max(errs) <= 1e-4
