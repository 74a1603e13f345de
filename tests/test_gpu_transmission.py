"""Full-length transmission runs at the two single-grid workload shapes against the FP64 oracle on
the same inputs (the oracle runs on the host cores: about 15 s and 85 s):

* the 240 x 120 x 100 Si strip waveguide (library 120 x 100 x 240): TE0 of the mode solver as the
  pulsed source at plane 20, forward overlaps at planes 25 and 220, 16 wavelengths 1.50-1.60 um,
  4000 steps through the device-resident run API;
* the 620 x 200 x 100 linear taper (library 200 x 100 x 620): TE0 of each port, forward run of
  6000 steps, overlap at the launch and output ports at 1.55 um.

Gate: max over wavelengths of |t_gpu - t_fp64| <= 1e-5 for the complex mode transmission
t = a_out / a_in (FP32 fields, FP64 monitor accumulation on the device)."""
import ctypes as C
import importlib
import math
import sys

import numpy as np
import pytest

from oracle import taper_oracle

pytestmark = pytest.mark.gpu

fd = importlib.import_module("3d_fdtd_amd")
design = importlib.import_module("3d_fdtd_amd.design")
WL = importlib.import_module("3d_fdtd_amd.workloads")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_strip_waveguide_transmission_vs_fp64_oracle():
    steps = 4000
    say("strip waveguide", WL.WG_I, WL.WG_J, WL.WG_K, steps, "steps vs FP64 oracle")
    L = fd.lib()
    g, dt, md, mons, omegas = WL.strip_waveguide(steps)
    fd.check(L.fdtd3d_ext_run(g.tf, dt, 0, steps, 0.0), "run")
    acc = []
    for m in (mons[0], mons[2]):
        cum = (fd.Complex * len(omegas))()
        fd.check(L.fdtd3d_ext_run_overlap(g.tf, m, cum, None), "overlap")
        acc.append(np.array([complex(c.re, c.im) for c in cum]))
    t_gpu = acc[1] / acc[0]
    t_ref = taper_oracle.transmission(np.array(g.a["ep"]), WL.WG_I, WL.WG_J, WL.WG_K, WL.DISC, dt, WL.WG_NPML,
                                      fd.mode_fields(md).copy(), 60, 50, g.amp, g.phase, omegas, 20,
                                      (25, WL.WG_K - 20))
    L.mode_free(md)
    g.free()
    err = np.abs(t_gpu - t_ref).max()
    say("|t|^2", np.round(np.abs(t_ref) ** 2, 4), "max |dt|", err)
    assert np.abs(t_ref).min() > 0.9
    assert err <= 1e-5


def test_taper_forward_transmission_vs_fp64_oracle():
    steps = 6000
    fam = design.TaperFamily(disc=WL.DISC, I=200, J=100, core_cells=11, npml=10, port=48, steps=steps,
                             lam=(1.55, 1.55), nfreq=1)
    p = np.concatenate([np.linspace(0.5, 3.0, 8), [10.0]])
    ep, K = fam.permittivity(p)
    say("taper", fam.I, fam.J, K, steps, "steps vs FP64 oracle")
    w0 = 2 * math.pi / 1.55
    ks, km = fam.npml + 1, K - fam.npml - 2
    m_in, mi_in, mj_in, _ = WL.port_mode(fam, p, ks, w0)
    m_out, mi_out, mj_out, _ = WL.port_mode(fam, p, km, w0)
    om = np.array([w0])
    g = fd.DevGrid(fam.I, fam.J, K, fam.disc, ep)
    g.cpml((fam.npml,) * 6, 3, 1e-8, 1.0, 0.05)
    ci, cj = fam.I // 2, fam.J // 2
    g.add_source(ci, cj, ks, m_in, mi_in, mj_in, fam.disc, fam.amp, fam.phase)
    a = g.add_overlap(ci, cj, ks + 3, m_in, mi_in, mj_in, fam.disc, om)
    b = g.add_overlap(ci, cj, km, m_out, mi_out, mj_out, fam.disc, om)
    g.run(fam.dt, 0, steps)
    t_gpu = g.overlap(b, 1)[0][0] / g.overlap(a, 1)[0][0]
    g.free()
    t_ref = taper_oracle.transmission(ep, fam.I, fam.J, K, fam.disc, fam.dt, fam.npml, m_in, mi_in, mj_in, fam.amp,
                                      fam.phase, om, ks, (ks + 3, km), out_mode=(m_out, mi_out, mj_out))[0]
    say("t gpu", t_gpu, "t fp64", t_ref, "|dt|", abs(t_gpu - t_ref))
    assert abs(t_ref) > 0.5
    assert abs(t_gpu - t_ref) <= 1e-5
