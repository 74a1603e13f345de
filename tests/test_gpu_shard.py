"""Z-slab sharding of one grid, shown on one GPU: two shard grids whose boundary planes hold the
neighbour's planes (fdtd_dev_copy_planes after each half step: Ex, Ey down after the E update,
Hx, Hy up after the H update; the CPML z slabs on the end shards) advance exactly like the
undivided grid (bit-identical fields: every cell sees the same inputs and the same kernel)."""
import ctypes as C
import importlib
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

fd = importlib.import_module("3d_fdtd_amd")
L = fd.lib


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_two_slab_shards_with_halo_planes_equal_one_grid():
    I, J, K, disc = 30, 24, 50, 0.02
    dt = 0.99 * disc / np.sqrt(3)
    cut = 25  # shard A owns global planes 1..24, shard B 25..48
    say("two z-slab shards of a", I, J, K, "grid, cut at plane", cut)
    rng = np.random.default_rng(21)
    ep = rng.uniform(1, 12, (K, J, I, 3)).astype(np.float32)
    f0 = rng.uniform(-1, 1, (K, J, I, 6)).astype(np.float32).astype(np.float64)
    grading = (3, 1e-6, 1.5, 0.1)

    def step(g):
        fd.dev_check(L().fdtd_dev_update_E_cpml(g.d, dt), "E")
        fd.dev_check(L().fdtd_dev_update_H_cpml(g.d, dt), "H")

    g = fd.DevGrid(I, J, K, disc, ep.reshape(-1))
    g.cpml((5, 5, 4, 4, 6, 6), *grading)
    g.set_fields(f0.reshape(-1))
    for _ in range(40):
        step(g)
    ref = g.fields().reshape(K, J, I, 6)
    g.free()

    a = fd.DevGrid(I, J, cut + 1, disc, ep[:cut + 1].reshape(-1))     # global 0..cut (cut = ghost)
    b = fd.DevGrid(I, J, K - cut + 1, disc, ep[cut - 1:].reshape(-1))  # global cut-1 (ghost)..K-1
    a.cpml((5, 5, 4, 4, 6, 0), *grading)
    b.cpml((5, 5, 4, 4, 0, 6), *grading)
    a.set_fields(f0[:cut + 1].reshape(-1))
    b.set_fields(f0[cut - 1:].reshape(-1))
    EXY, HXY = 0b000011, 0b011000
    for _ in range(40):
        fd.dev_check(L().fdtd_dev_update_E_cpml(a.d, dt), "E a")
        fd.dev_check(L().fdtd_dev_update_E_cpml(b.d, dt), "E b")
        fd.dev_check(L().fdtd_dev_copy_planes(a.d, cut, b.d, 1, EXY), "halo E")
        fd.dev_check(L().fdtd_dev_update_H_cpml(a.d, dt), "H a")
        fd.dev_check(L().fdtd_dev_update_H_cpml(b.d, dt), "H b")
        fd.dev_check(L().fdtd_dev_copy_planes(b.d, 0, a.d, cut - 1, HXY), "halo H")
    fa = a.fields().reshape(cut + 1, J, I, 6)
    fb = b.fields().reshape(K - cut + 1, J, I, 6)
    got = np.concatenate([fa[:cut], fb[1:]])
    diff = np.abs(got - ref).max()
    say("max |sharded - single grid| over all cells:", diff)
    assert np.abs(ref).max() > 0
    assert diff == 0.0
    # the argument checks of the copy
    assert L().fdtd_dev_copy_planes(a.d, cut + 1, b.d, 1, EXY) == -2
    assert L().fdtd_dev_copy_planes(a.d, cut, b.d, 1, 0) == -2
    a.free()
    b.free()
