"""TEST INFRASTRUCTURE ONLY -- FP64 CPU run of one taper design through the oracle (fdtd_oracle.c):
the checker of the device's batched taper transmissions, and the CPU baseline of the design
workload.  Nothing in the product imports it.

One run: for each step n, the pulsed mode source on plane src_k (orc_mode_source, phase[n],
amp[n]; timefield3d.c:2977-3026 semantics), one CPML step (orc_step_cpml), then the overlap of
every monitor plane (orc_mode_overlap, timefield3d.c:3176-3239), DFT'd in FP64 with e^{i w n dt}.
The transmission is the ratio of the last monitor's running DFT to the first's.
"""
import numpy as np

from oracle import oracle_py as O


def run(ep, I, J, K, disc, dt, npml, mode, mi, mj, amp, phase, omegas, src_k, mon_ks,
        cpml=dict(order=3, R0=1e-8, kappa_max=1.0, alpha_max=0.05), out_mode=None, nsteps=None):
    """The first nsteps steps (default len(amp)) of the run: returns (fields [k][j][i][6] FP64,
    [running DFT of each monitor over omegas]).  Inputs as for transmission()."""
    g = O.Grid(I, J, K, disc)
    g.ep[:] = ep
    cp = O.Cpml(g, (npml,) * 6, dt, **cpml)
    mode = np.array(mode, np.float64)
    omegas = np.asarray(omegas, np.float64)
    ci, cj = I // 2, J // 2
    acc = [np.zeros(len(omegas), complex) for _ in mon_ks]
    cum, inst = np.zeros(2), np.zeros(2)
    L = O.lib()
    mons = [(mode, mi, mj)] * len(mon_ks)
    if out_mode is not None:
        mons[-1] = (np.array(out_mode[0], np.float64), out_mode[1], out_mode[2])
    for n in range(len(amp) if nsteps is None else nsteps):
        L.orc_mode_source(g.ptr, ci, cj, src_k, mi, mj, disc, O.dptr(mode), float(phase[n]), float(amp[n]), 0)
        O.step_cpml(g, cp, dt)
        for q, kz in enumerate(mon_ks):
            m, a, b = mons[q]
            L.orc_mode_overlap(g.ptr, ci, cj, kz, a, b, disc, O.dptr(m), 0.0, 0, O.dptr(cum), O.dptr(inst))
            acc[q] += np.exp(1j * omegas * n * dt) * complex(inst[0], inst[1])
    return np.array(g.f), acc


def transmission(ep, I, J, K, disc, dt, npml, mode, mi, mj, amp, phase, omegas, src_k, mon_ks,
                 cpml=dict(order=3, R0=1e-8, kappa_max=1.0, alpha_max=0.05), out_mode=None):
    """Complex t(omega) = a(mon_ks[-1]) / a(mon_ks[0]) of the design with permittivity ep
    ([K][J][I][3] float32, flattened); mode is the power-normalised interleaved complex
    [mj][mi][6] window centred on (I/2, J/2), used by the source and every monitor unless
    out_mode = (mode, mi, mj) gives the last monitor its own (a port of another width)."""
    _, acc = run(ep, I, J, K, disc, dt, npml, mode, mi, mj, amp, phase, omegas, src_k, mon_ks, cpml, out_mode)
    return acc[-1] / acc[0]
