"""The benchmark configurations as device-API set-ups, shared by bench.py and the GPU tests (so a
parity test runs exactly the workload the benchmark times).

All grids are stored in the library's own axes (the reference's: k = propagation, j = vertical,
timefield3d.c:1368).  A configuration stated with x as the propagation axis maps (x, y, z) ->
library (k, i, j): a 240 x 120 x 100 configuration grid is the library grid I = 120, J = 100,
K = 240.

* strip_waveguide: straight Si strip waveguide, 240 x 120 x 100 cells of 20 nm, 220 x 500 nm core
  (n = 3.48) in SiO2 (n = 1.44), 10-cell CPML, TE0 of the mode solver as a pulsed soft source at
  x = 20, FP64 overlap monitors (launched / reflected at x = 25, transmitted at x = 220) with a
  running DFT at 16 wavelengths 1.50-1.60 um.
* tapers: 64 Si tapers (seeded): length L ~ U[3, 15] um, 8 spline control-point widths
  ~ U[0.4, 3.2] um, 200 x 100 cross-section, 8-cell CPML, a mode source after the input PML and an
  FP64 overlap monitor with a running DFT at 16 wavelengths before the output PML; each taper is
  its own grid of uneven length (170-770 cells), rasterised on the device.
"""
import ctypes as C
import importlib
import math
import sys

import numpy as np

fd = importlib.import_module("3d_fdtd_amd")

DISC = 0.02
LAM0, DLAM = 1.55, 0.2

# ---- strip waveguide ----
WG_I, WG_J, WG_K = 120, 100, 240
WG_NPML = 10


def pulse(steps_total, dt, lam0=LAM0, dlam=DLAM):
    """Gaussian-modulated carrier: amplitude[n], phase[n] of the soft source at step n (the
    envelope peaks 4 sigma after the start, sigma = lam0^2 / (2 pi dlam) in time units)."""
    w0 = 2 * math.pi / lam0
    sig_t = lam0 * lam0 / (2 * math.pi * dlam)
    t = np.arange(steps_total) * dt
    return np.exp(-(t - 4 * sig_t) ** 2 / (2 * sig_t ** 2)), w0 * t


def band_omegas(n=16, lo=1.50, hi=1.60):
    return 2 * math.pi / np.linspace(lo, hi, n)


def te0_mode(w0, mi=60, mj=50, log=True):
    """Fundamental (TE0) mode of the strip's cross-section from the library's mode solver
    (mode_solve_beta_squared): a mi x mj window centred on the 25 x 11-cell Si core, zero walls.
    Returns the library mode struct (the caller frees it with mode_free)."""
    ep = np.full((mj, mi, 3), 1.44 ** 2, complex)
    x0, y0 = mi // 2 - 12, mj // 2 - 5
    ep[y0:y0 + 11, x0:x0 + 25, :] = 3.48 ** 2
    beta, md = fd.solve_mode(ep, DISC, w0, tol=1e-9, size=20, confinement=2)
    if log:
        print(f"workload: TE0 mode n_eff {beta.real / w0:.4f}", file=sys.stderr, flush=True)
    return md


def strip_waveguide(steps_total):
    """The strip waveguide on the drop-in surface with the device-resident run API.  Returns
    (grid, dt, mode struct, monitor ids [launched, reflected, transmitted], omegas); the grid
    carries the source schedule as g.amp, g.phase."""
    L = fd.lib()
    dt = 0.99 * DISC / math.sqrt(3)
    g = fd.Grid(WG_I, WG_J, WG_K, DISC)
    L.timefield3d_set_background(g.tf, 1.44 ** 2, 1.0)
    ci, cj = WG_I // 2, WG_J // 2
    L.timefield3d_set_prism_permittivity(g.tf, 3.48 ** 2, ci - 12, ci + 13, cj - 5, cj + 6, 0, WG_K)
    npml = (C.c_int * 6)(*([WG_NPML] * 6))
    fd.check(L.fdtd3d_ext_set_cpml(g.tf, npml, 3, 1e-8, 1.0, 0.05), "set_cpml")
    w0 = 2 * math.pi / LAM0
    md = te0_mode(w0)
    amp, phase = pulse(steps_total, dt)
    dptr = C.POINTER(C.c_double)
    src = L.fdtd3d_ext_run_add_mode_source(g.tf, ci, cj, 20, md, 0, 0, steps_total, amp.ctypes.data_as(dptr),
                                           phase.ctypes.data_as(dptr))
    fd.check(0 if src >= 0 else src, "add source")
    g.amp, g.phase = amp, phase
    omegas = band_omegas()
    mons = []
    for kz, rev in ((25, 0), (25, 1), (WG_K - 20, 0)):
        m = L.fdtd3d_ext_run_add_overlap(g.tf, ci, cj, kz, md, rev, len(omegas), omegas.ctypes.data_as(dptr))
        fd.check(0 if m >= 0 else m, "add overlap")
        mons.append(m)
    return g, dt, md, mons, omegas


# ---- batched tapers ----
TAPER_SEED = 20261203
TAPER_I, TAPER_J = 200, 100  # lateral x vertical cross-section (library i, j), 20 nm cells
TAPER_NPML = 8
TAPER_CORE = 11              # 220 nm core height in cells
TAPER_MODE = (60, 50)        # source / monitor mode window (cells)
EPS_CLAD, EPS_CORE = np.float32(1.44 ** 2), np.float32(3.48 ** 2)


def taper_geometry(n, seed=TAPER_SEED):
    """n tapers: length L ~ U[3, 15] um, 8 control-point widths ~ U[0.4, 3.2] um (seeded)."""
    rng = np.random.default_rng(seed)
    return rng.uniform(3.0, 15.0, n), rng.uniform(0.4, 3.2, (n, 8))


def taper_length_cells(L):
    return int(round(L / DISC)) + 20


def taper_core(L, widths):
    """Core mask [k][i] of one taper: the core is |x| < w(k)/2 about the grid centre (cell centres
    at i + 0.5 - I/2), w(k) the cubic spline through the 8 control points, held constant beyond
    the end points (the strip ports)."""
    from scipy.interpolate import CubicSpline

    K = taper_length_cells(L)
    k = np.arange(K)
    xs = np.linspace(10, K - 11, len(widths))
    w = CubicSpline(xs, widths)(np.clip(k, xs[0], xs[-1])) / DISC  # width in cells
    i = np.arange(TAPER_I) + 0.5 - TAPER_I / 2
    return np.abs(i)[None, :] < (w[:, None] / 2), K


def taper_rows():
    j0 = TAPER_J // 2 - TAPER_CORE // 2
    return j0, j0 + TAPER_CORE


def taper_permittivity(L, widths):
    """eps of one taper grid as the host raster [k][j][i][3] (float32, flattened): SiO2 cladding,
    the Si core on the TAPER_CORE rows around the centre where taper_core holds."""
    core, K = taper_core(L, widths)
    ep = np.full((K, TAPER_J, TAPER_I, 3), EPS_CLAD, np.float32)
    j0, j1 = taper_rows()
    ep[:, j0:j1, :, :][np.broadcast_to(core[:, None, :, None], (K, j1 - j0, TAPER_I, 3))] = EPS_CORE
    return ep.reshape(-1), K


def taper_extruded(L, widths):
    """The same geometry as taper_permittivity in the device raster's form
    (fdtd_dev_set_materials_extruded): table [[clad]*3, [core]*3], one id per (k, i) for the core
    rows, background id 0 elsewhere."""
    core, K = taper_core(L, widths)
    table = np.array([[EPS_CLAD] * 3, [EPS_CORE] * 3], np.float32)
    j0, j1 = taper_rows()
    return (table, core.astype(np.uint16), j0, j1, 0), K


def taper_mode():
    """The tapers' source / monitor mode: a Gaussian-profile mode of the 60 x 50 window in the
    cladding (mode_gaussian), power-normalised; as interleaved complex doubles."""
    L = fd.lib()
    w0 = 2 * math.pi / LAM0
    mi, mj = TAPER_MODE
    md = L.mode_gaussian(mi, mj, DISC, fd.Complex(1.44 ** 2, 0), fd.Complex(1.0, 0), w0, 8.0, 0)
    mode = fd.normalised_mode(mi, mj, DISC, fd.mode_fields(md).copy())
    L.mode_free(md)
    return mode


def taper_grid(L, widths, mode, amp, phase, omegas, host_raster=False):
    """One taper as a device grid with its CPML, mode source (plane npml + 1) and transmitted
    overlap monitor (plane K - npml - 2); the monitor id is g.mon."""
    if host_raster:
        ep, K = taper_permittivity(L, widths)
        g = fd.DevGrid(TAPER_I, TAPER_J, K, DISC, ep)
    else:
        ex, K = taper_extruded(L, widths)
        g = fd.DevGrid(TAPER_I, TAPER_J, K, DISC, None, extruded=ex)
    g.cpml((TAPER_NPML,) * 6, 3, 1e-8, 1.0, 0.05)
    mi, mj = TAPER_MODE
    g.add_source(TAPER_I // 2, TAPER_J // 2, TAPER_NPML + 1, mode, mi, mj, DISC, amp, phase)
    g.mon = g.add_overlap(TAPER_I // 2, TAPER_J // 2, K - TAPER_NPML - 2, mode, mi, mj, DISC, omegas)
    return g


def tapers(n_tapers, steps_total, select=None):
    """The batched taper set: (grids, dt, omegas, lengths, widths, mode, amp, phase).  select:
    indices of the seeded set to build (default all n_tapers)."""
    lengths, widths = taper_geometry(n_tapers)
    dt = 0.99 * DISC / math.sqrt(3)
    amp, phase = pulse(steps_total, dt)
    omegas = band_omegas()
    mode = taper_mode()
    idx = range(n_tapers) if select is None else select
    grids = []
    try:
        for q in idx:
            grids.append(taper_grid(lengths[q], widths[q], mode, amp, phase, omegas))
    except Exception:
        for g in grids:
            g.free()
        raise
    return grids, dt, omegas, lengths, widths, mode, amp, phase


def port_mode(fam, p, k, w0, margin=12):
    """TE0 of the cross-section at plane k of design p (the mode solver on a window inside the
    PML), as the power-normalised interleaved array the run API takes, its window size and n_eff."""
    ep, K = fam.permittivity(p)
    sec = ep.reshape(K, fam.J, fam.I, 3)[k].astype(np.complex128)
    mi, mj = fam.I - 2 * margin, fam.J - 2 * margin
    i0, j0 = fam.I // 2 - mi // 2, fam.J // 2 - mj // 2
    beta, md = fd.solve_mode(sec[j0:j0 + mj, i0:i0 + mi], fam.disc, w0, tol=1e-9, size=20, confinement=2)
    arr = fd.normalised_mode(mi, mj, fam.disc, fd.mode_fields(md).copy())
    fd.lib().mode_free(md)
    return arr, mi, mj, beta / w0
