"""Grids of 2^30 cells and more: the split E/H kernels address them with 64-bit cell offsets (every
smaller grid takes 32-bit offsets and a compile-time z chunk), and the Mur faces, shell snapshot,
source planes and CPML slabs index them with 64-bit arithmetic.

A 1024 x 1024 x 1100 grid (1.15 G cells, 27.7 GB of FP32 fields) and a 1024 x 1024 x 16 grid get
the same random plane source, at z plane 1060 of the large grid (cell offsets past 2^30, byte
offsets past 4 GiB) and at plane 8 of the small one, and advance the same number of steps from
zero fields.  A step moves a disturbance at most one plane up or down, so after 6 steps both grids
differ from zero only within planes 2..13 of the source window and never reach a z face; their
lateral faces (Mur faces, or CPML slabs in front of PEC walls) are the same, and the source covers
the whole plane so that those faces take part.  The 16 planes around the large grid's source,
pulled into a third small grid with fdtd_dev_copy_planes, must therefore equal the small grid's
fields: the same cells see the same inputs and the same arithmetic, only the offsets differ (the
small grid's unrolled two-plane variant may round differently from the large grid's looped one,
hence a 1e-6 relative L2 bound rather than bit equality)."""
import ctypes as C
import importlib
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

fd = importlib.import_module("3d_fdtd_amd")
L = fd.lib

I, J = 1024, 1024
K_BIG, K_SMALL = 1100, 16
KB, KS = 1060, 8  # source planes
NSTEPS = 6
DISC = 0.02


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_grid(K):
    ids = np.zeros((K, I), np.uint16)
    table = np.array([[2.25, 2.25, 2.25]], np.float32)
    return fd.DevGrid(I, J, K, DISC, None, extruded=(table, ids, 0, 0, 0))


@pytest.mark.parametrize("boundary", ["mur", "cpml"])
def test_grid_past_2_30_cells_equals_small_grid(boundary):
    assert I * J * K_BIG >= 1 << 30 and I * J * K_SMALL < 1 << 30
    dt = 0.99 * DISC / np.sqrt(3)
    rng = np.random.default_rng(1060 if boundary == "mur" else 1061)
    plane = rng.uniform(-1, 1, 6 * I * J).astype(np.float32)
    fp = C.POINTER(C.c_float)
    say("wide offsets:", boundary, "boundary,", I, J, K_BIG, "grid,", NSTEPS, "steps")
    grids = []
    try:
        for K, ks in ((K_BIG, KB), (K_SMALL, KS)):
            g = make_grid(K)
            grids.append(g)
            if boundary == "cpml":
                g.cpml((8, 8, 8, 8, 0, 0), 3, 1e-6, 1.5, 0.1)
            fd.dev_check(L().fdtd_dev_set_source_plane(g.d, ks, plane.ctypes.data_as(fp), None), "source plane")
            for _ in range(NSTEPS):
                fd.dev_check(L().fdtd_dev_step(g.d, dt, 0), "step")
            fd.dev_check(L().fdtd_dev_synchronize(g.d), "sync")
            say("  grid", K, "planes stepped")
        big, small = grids
        window = make_grid(K_SMALL)
        grids.append(window)
        for k in range(K_SMALL):
            fd.dev_check(L().fdtd_dev_copy_planes(window.d, k, big.d, KB - KS + k, 0x3F), "copy planes")
        fd.dev_check(L().fdtd_dev_synchronize(window.d), "sync")
        want = small.fields().reshape(K_SMALL, J, I, 6)
        got = window.fields().reshape(K_SMALL, J, I, 6)
    finally:
        for g in grids:
            g.free()
    live = [k for k in range(K_SMALL) if np.any(want[k] != 0)]
    say("  nonzero planes of the small grid:", live)
    assert live and min(live) >= 2 and max(live) <= 13
    assert np.any(want[:, :, :8] != 0) and np.any(want[:, :8] != 0)  # the lateral faces took part
    diff = float(np.abs(got - want).max())
    err = float(np.linalg.norm(got - want) / np.linalg.norm(want))
    say("  max |diff|", diff, "rel L2", err)
    assert err <= 1e-6
