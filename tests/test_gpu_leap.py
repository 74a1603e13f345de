"""GPU tests of the single-pass leap launches (fdtd_dev_set_step_kernel(d, -2, kc)): a device-resident
run issued as E(s0), then H(s) with E(s+1) per launch, then H(s0+n-1) must equal the same run on
the split E / H kernels, fields and FP64 overlap monitors to relative 1e-6 (the same updates in the
same order, so FP32 rounding at most).

Cases cover z chunks from one plane to longer than the grid (1 plane: every launch recomputes
the plane below each chunk), grids whose x and y tiles and z chunks do not divide evenly,
uneven CPML slabs on the six faces, two pulsed run sources on different planes, two overlap
monitors (one reversed), seeded random initial fields in every cell (so the PML memories and the
halo recomputation carry real values), several run calls in a row (n0 continuation) and a run
call of a single step."""
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


def _make(I, J, K, npml, seed, kernel, shape):
    rng = np.random.default_rng(seed)
    ep = np.full((K, J, I, 3), 2.0, np.float32)
    ep[K // 4:3 * K // 4, J // 3:J // 2, I // 3:2 * I // 3, :] = 6.5
    ep[..., 1] += (rng.integers(0, 1000, (K, J, I)) / 2000).astype(np.float32)  # > 256 materials: 16-bit ids
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
    L = fd.lib()
    if kernel == "leap":
        fd.dev_check(L.fdtd_dev_set_step_kernel(g.d, -2, shape), "step kernel")
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
    (130, 31, 37, (6, 6, 6, 6, 6, 6)),   # three x strips of 43 columns; 29 rows
]
SHAPES = [1, 2, 3, 5, 8, 64, 0]  # z planes per chunk (0: automatic)


@pytest.mark.parametrize("grid", GRIDS, ids=["40x30x50", "130x31x37"])
@pytest.mark.parametrize("shape", SHAPES, ids=[f"kc{k}" for k in SHAPES])
def test_leap_run_equals_split_run(grid, shape):
    I, J, K, npml = grid
    calls = (24, 1, NSTEPS - 25)
    say(f"leap vs split: grid {I}x{J}x{K}, CPML {npml}, z chunk {shape}, run calls {calls}")
    gs, ms, nf = _make(I, J, K, npml, 11, "split", None)
    gl, ml, _ = _make(I, J, K, npml, 11, "leap", shape)
    try:
        _run(gs, calls)
        _run(gl, calls)
        mon_s, f_s = _results(gs, ms, nf)
        mon_l, f_l = _results(gl, ml, nf)
    finally:
        gs.free()
        gl.free()
    ef, em = rel(f_l, f_s), rel(mon_l, mon_s)
    say(f"  fields rel {ef:.2e}, monitors rel {em:.2e}")
    assert ef <= 1e-6
    assert em <= 1e-6
