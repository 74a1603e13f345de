"""GPU tests of the single-pass full step (fdtd_dev_set_step_kernel(d, -3, kc)): launches per time
step that form the new E of a z plane and then the new H of the plane below, reading one field
buffer and writing the other.  A device-resident run on it must equal the same run on the split
E / H kernels, fields and FP64 overlap monitors to relative 1e-6 (the same per-cell arithmetic in
the same order, so FP32 rounding at most; the monitor partials are summed over other cell groups).

Cases cover z chunks from one plane to longer than the grid (4: every load of the chunk issued up
front, with short last chunks on two grids; others: the rotating prefetch), grids whose x and y
tiles and z chunks do not divide evenly, uneven CPML slabs on the six faces (so every slab class
of blocks runs: clear of the slabs, x slab, y slab, both, z-end chunks; each with and without
source planes), 1- and 2-byte material ids, isotropic lossless media (one coefficient per
material) and anisotropic ones (the general coefficient table), two pulsed run sources on
different planes, two overlap monitors (one reversed), seeded random initial fields in every cell
(so the PML memories and the tile-edge recomputation carry real values), several run calls in a
row and a run call of a single step."""
import importlib
import sys

import numpy as np
import pytest

from tests import gradient_case as GC

pytestmark = pytest.mark.gpu

fd = importlib.import_module("3d_fdtd_amd")

NSTEPS = 61


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _make(I, J, K, npml, seed, rows, kc, media):
    rng = np.random.default_rng(seed)
    ep = np.full((K, J, I, 3), 2.0, np.float32)
    ep[K // 4:3 * K // 4, J // 3:J // 2, I // 3:2 * I // 3, :] = 6.5
    if media == "aniso16":  # > 256 anisotropic materials: 16-bit ids, the general coefficient table
        ep[..., 1] += (rng.integers(0, 1000, (K, J, I)) / 2000).astype(np.float32)
    elif media == "iso16":  # > 256 isotropic materials: 16-bit ids, the isotropic lossless table
        ep += (rng.integers(0, 1000, (K, J, I, 1)) / 2000).astype(np.float32)
    g = fd.DevGrid(I, J, K, GC.DISC, ep.reshape(-1))
    g.cpml(npml)
    g.set_fields(rng.uniform(-1, 1, 6 * I * J * K))
    mode = fd.normalised_mode(GC.MI, GC.MJ, GC.DISC, GC.mode_array())
    n = np.arange(NSTEPS + 5)
    amp = np.exp(-((n - 30.0) / 12.0) ** 2 / 2)
    g.add_source(I // 2, J // 2, npml[4] + 2, mode, GC.MI, GC.MJ, GC.DISC, amp, GC.OMEGA * n * GC.DT)
    g.add_source(I // 2 + 1, J // 2, K - npml[5] - 5, mode, GC.MI, GC.MJ, GC.DISC, 0.5 * amp, GC.OMEGA * n * GC.DT + 1)
    oms = GC.OMEGA * np.array([0.9, 1.0, 1.1])
    mons = [g.add_overlap(I // 2, J // 2, K - npml[5] - 3, mode, GC.MI, GC.MJ, GC.DISC, oms),
            g.add_overlap(I // 2, J // 2, npml[4] + 4, mode, GC.MI, GC.MJ, GC.DISC, oms, reverse=1)]
    if rows is not None:
        fd.dev_check(fd.lib().fdtd_dev_set_step_kernel(g.d, rows, kc), "step kernel")
    return g, mons, len(oms)


def _run(g, calls):
    n0 = 0
    for n in calls:
        g.run(GC.DT, n0, n)
        n0 += n


def _results(g, mons, nf):
    out = []
    for m in mons:
        cum, last = g.overlap(m, nf)
        out.extend(list(cum) + [last])
    return np.array(out), g.fields()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


GRIDS = [
    # I, J, K, CPML per face (x-lo, x-hi, y-lo, y-hi, z-lo, z-hi)
    (40, 30, 50, (5, 7, 4, 6, 3, 8)),
    (130, 31, 37, (6, 6, 6, 6, 6, 6)),
    (100, 60, 24, (8, 8, 8, 8, 8, 8)),  # interior tiles without CPML memories
]
CASES = [(-3, 0), (-3, 1), (-3, 3), (-3, 2), (-3, 4), (-3, 5), (-3, 8), (-3, 64)]


@pytest.mark.parametrize("grid", GRIDS, ids=["40x30x50", "130x31x37", "100x60x24"])
@pytest.mark.parametrize("rows,kc", CASES, ids=[f"kc{k}" for r, k in CASES])
@pytest.mark.parametrize("media", ["aniso16", "iso16", "iso8"])
def test_onepass_run_equals_split_run(grid, rows, kc, media):
    I, J, K, npml = grid
    calls = (24, 1, NSTEPS - 25)
    say(f"single-pass vs split: grid {I}x{J}x{K}, CPML {npml}, rows {rows}, z chunk {kc}, media {media}, "
        f"run calls {calls}")
    gs, ms, nf = _make(I, J, K, npml, 11, None, None, media)
    go, mo, _ = _make(I, J, K, npml, 11, rows, kc, media)
    try:
        _run(gs, calls)
        _run(go, calls)
        mon_s, f_s = _results(gs, ms, nf)
        mon_o, f_o = _results(go, mo, nf)
    finally:
        gs.free()
        go.free()
    ef, em = rel(f_o, f_s), rel(mon_o, mon_s)
    say(f"  fields rel {ef:.2e}, monitors rel {em:.2e}")
    assert ef <= 1e-6
    assert em <= 1e-6
