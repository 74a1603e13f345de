"""GPU tests of batched multi-geometry runs (fdtd_dev_batch_*): grids of uneven length advanced in
lockstep by persistent kernels over a descriptor table, steps replayed from a captured hipGraph,
z-chunk depth picked by the autotuner.  The batched result must equal each grid run on its own
through the single-grid path (same arithmetic, so to FP32 rounding: relative 1e-6) for fields and
FP64 overlap monitors, including across several run calls and with a fixed chunk depth."""
import importlib
import sys

import numpy as np
import pytest

from tests import gradient_case as GC

pytestmark = pytest.mark.gpu

fd = importlib.import_module("3d_fdtd_amd")

KS = (30, 47, 64)
NSTEPS = 170  # more than the 10 x 5 steps the shape autotuner times, so graph segments run too


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _schedule():
    n = np.arange(NSTEPS)
    return np.exp(-((n - 60.0) / 20.0) ** 2 / 2), GC.OMEGA * n * GC.DT


def _make(K, w):
    I = J = 20
    ep = np.full((K, J, I, 3), 2.0, np.float32)
    ep[:, 8:12, 10 - w:10 + w, :] = 6.0
    g = fd.DevGrid(I, J, K, GC.DISC, ep.reshape(-1))
    g.cpml((5, 5, 5, 5, 5, 5))
    mode = fd.normalised_mode(GC.MI, GC.MJ, GC.DISC, GC.mode_array())
    amp, phase = _schedule()
    g.add_source(I // 2, J // 2, 7, mode, GC.MI, GC.MJ, GC.DISC, amp, phase)
    oms = GC.OMEGA * np.array([0.9, 1.0, 1.1])
    mons = [g.add_overlap(I // 2, J // 2, K - 8, mode, GC.MI, GC.MJ, GC.DISC, oms),
            g.add_overlap(I // 2, J // 2, 10, mode, GC.MI, GC.MJ, GC.DISC, oms, reverse=1)]
    return g, mons, len(oms)


def _results(g, mons, nf):
    out = []
    for m in mons:
        cum, last = g.overlap(m, nf)
        out.extend(list(cum) + [last])
    return np.array(out), g.fields()


def _single():
    res = []
    for q, K in enumerate(KS):
        g, mons, nf = _make(K, 1 + q)
        g.run(GC.DT, 0, NSTEPS)
        res.append(_results(g, mons, nf))
        g.free()
    return res


@pytest.fixture(scope="module")
def single_results():
    say("batch: single-grid reference runs")
    return _single()


@pytest.mark.parametrize("calls,shape", [((0, NSTEPS), None), ((0, 100, NSTEPS), None), ((0, 37, NSTEPS), (3, True, True)),
                                         ((0, NSTEPS), (5, False, False)), ((0, 41, NSTEPS), (2, False, True))])
def test_batch_matches_single_grids(single_results, calls, shape):
    say("batch: calls", calls, "shape", shape)
    built = [_make(K, 1 + q) for q, K in enumerate(KS)]
    b = fd.Batch([g for g, _, _ in built])
    if shape:
        b.set_shape(*shape)
    for a, c in zip(calls[:-1], calls[1:]):
        b.run(GC.DT, a, c - a)
    info = b.info()
    say("batch info", info)
    assert info["cells"] == sum(20 * 20 * K for K in KS)
    if shape:
        assert info["shape"] == fd.Batch.shape_name(*shape)
    else:
        assert len(info["tuned"]) == 11 and info["shape"] in info["tuned"]  # every tuner candidate timed
    for (g, mons, nf), (mon_ref, f_ref) in zip(built, single_results):
        mon, f = _results(g, mons, nf)
        em = np.abs(mon - mon_ref).max() / np.abs(mon_ref).max()
        ef = np.linalg.norm(f - f_ref) / np.linalg.norm(f_ref)
        say("grid K", g.K, "monitor max rel", em, "field rel L2", ef)
        assert np.abs(mon_ref).max() > 0
        assert em <= 1e-6 and ef <= 1e-6
    b.free()
    for g, _, _ in built:
        g.free()


def test_batch_of_mixed_material_id_widths():
    """A grid with more than 256 distinct materials (2-byte ids) batched with grids of 1-byte ids:
    the batch widens the narrow grids' ids to one width, and every grid still equals its own
    single-grid run."""
    def make(q, K, graded):
        g, mons, nf = _make(K, 1 + q)
        if graded:  # 400 distinct permittivities in the core rows
            I = J = 20
            ep = np.full((K, J, I, 3), 2.0, np.float32)
            ep[:, 8:12, :, :] = (2.0 + np.arange(K * 4 * I).reshape(K, 4, I, 1) % 400 / 100.0).astype(np.float32)
            g.free()
            g = fd.DevGrid(I, J, K, GC.DISC, ep.reshape(-1))
            g.cpml((5, 5, 5, 5, 5, 5))
            mode = fd.normalised_mode(GC.MI, GC.MJ, GC.DISC, GC.mode_array())
            amp, phase = _schedule()
            g.add_source(I // 2, J // 2, 7, mode, GC.MI, GC.MJ, GC.DISC, amp, phase)
            oms = GC.OMEGA * np.array([0.9, 1.0, 1.1])
            mons = [g.add_overlap(I // 2, J // 2, K - 8, mode, GC.MI, GC.MJ, GC.DISC, oms),
                    g.add_overlap(I // 2, J // 2, 10, mode, GC.MI, GC.MJ, GC.DISC, oms, reverse=1)]
        return g, mons, nf

    spec = [(0, 30, False), (1, 47, True), (2, 64, False)]
    ref = []
    for q, K, graded in spec:
        g, mons, nf = make(q, K, graded)
        g.run(GC.DT, 0, NSTEPS)
        ref.append(_results(g, mons, nf))
        g.free()
    built = [make(q, K, graded) for q, K, graded in spec]
    b = fd.Batch([g for g, _, _ in built])
    b.run(GC.DT, 0, NSTEPS)
    for (g, mons, nf), (mon_s, f_s) in zip(built, ref):
        mon_b, f_b = _results(g, mons, nf)
        ef = float(np.linalg.norm(f_b - f_s) / np.linalg.norm(f_s))
        em = float(np.linalg.norm(mon_b - mon_s) / np.linalg.norm(mon_s))
        say("mixed ids: fields rel %.2e, monitors rel %.2e" % (ef, em))
        assert ef <= 1e-6 and em <= 1e-6
    b.free()
    for g, _, _ in built:
        g.free()


def test_batch_finish_with_uneven_monitor_sets():
    """Grids of one batch with two monitors of 3 wavelengths, none, and one of 5 wavelengths: the
    batched finish (one twiddle block per grid of its own row length, one table of every grid's
    monitors) gives each grid the monitors and fields of its own single-grid run, over two calls
    of which the second is the longer (the staging grows between them)."""
    def make(q, K):
        I = J = 20
        ep = np.full((K, J, I, 3), 2.0, np.float32)
        ep[:, 8:12, 9 - q:11 + q, :] = 6.0
        g = fd.DevGrid(I, J, K, GC.DISC, ep.reshape(-1))
        g.cpml((5, 5, 5, 5, 5, 5))
        mode = fd.normalised_mode(GC.MI, GC.MJ, GC.DISC, GC.mode_array())
        amp, phase = _schedule()
        g.add_source(I // 2, J // 2, 7, mode, GC.MI, GC.MJ, GC.DISC, amp, phase)
        if q == 0:
            oms = GC.OMEGA * np.array([0.9, 1.0, 1.1])
            mons = [g.add_overlap(I // 2, J // 2, K - 8, mode, GC.MI, GC.MJ, GC.DISC, oms),
                    g.add_overlap(I // 2, J // 2, 10, mode, GC.MI, GC.MJ, GC.DISC, oms, reverse=1)]
        elif q == 1:
            oms, mons = np.zeros(0), []
        else:
            oms = GC.OMEGA * np.array([0.8, 0.9, 1.0, 1.1, 1.2])
            mons = [g.add_overlap(I // 2, J // 2, K - 9, mode, GC.MI, GC.MJ, GC.DISC, oms)]
        return g, mons, len(oms)

    spec = [(0, 33), (1, 41), (2, 52)]
    ref = []
    for q, K in spec:
        g, mons, nf = make(q, K)
        g.run(GC.DT, 0, NSTEPS)
        ref.append(_results(g, mons, nf))
        g.free()
    built = [make(q, K) for q, K in spec]
    b = fd.Batch([g for g, _, _ in built])
    b.set_shape(2, True, False)
    b.run(GC.DT, 0, 61)
    b.run(GC.DT, 61, NSTEPS - 61)
    for (g, mons, nf), (mon_s, f_s) in zip(built, ref):
        mon_b, f_b = _results(g, mons, nf)
        ef = float(np.linalg.norm(f_b - f_s) / np.linalg.norm(f_s))
        say("uneven monitors: grid K %d, %d monitors, fields rel %.2e" % (g.K, len(mons), ef))
        assert ef <= 1e-6
        if mons:
            em = float(np.abs(mon_b - mon_s).max() / np.abs(mon_s).max())
            say("   monitors max rel %.2e" % em)
            assert np.abs(mon_s).max() > 0 and em <= 1e-6
    b.free()
    for g, _, _ in built:
        g.free()


def _extruded(q, K):
    """A 20 x 20 x K grid rasterised on the device from an extruded table: core rows 8..11 take
    ids[k][i] (core where |i - 10| < w(k)), every other cell the cladding."""
    I = J = 20
    table = np.array([[2.0] * 3, [6.0] * 3], np.float32)
    ids = np.zeros((K, I), np.uint16)
    for k in range(K):
        w = 1 + q + (k * 3) // K
        ids[k, 10 - w:10 + w] = 1
    g = fd.DevGrid(I, J, K, GC.DISC, None, extruded=(table, ids, 8, 12, 0))
    g.cpml((5, 5, 5, 5, 5, 5))
    mode = fd.normalised_mode(GC.MI, GC.MJ, GC.DISC, GC.mode_array())
    amp, phase = _schedule()
    g.add_source(I // 2, J // 2, 7, mode, GC.MI, GC.MJ, GC.DISC, amp, phase)
    oms = GC.OMEGA * np.array([0.9, 1.0, 1.1])
    mons = [g.add_overlap(I // 2, J // 2, K - 8, mode, GC.MI, GC.MJ, GC.DISC, oms)]
    return g, mons, len(oms)


@pytest.mark.parametrize("repaint", [False, True])
def test_batch_of_extruded_grids(repaint):
    """Grids rasterised from extruded tables: the batch takes their ids from the K x I tables
    (no id volume read) and equals each grid's single-grid run, which reads the volume.  With one
    grid repainted after the raster (a prism of another permittivity), that grid's table no longer
    describes its volume and is dropped, the batch falls back to the volumes, and it still equals
    the single-grid runs of the same repainted grids."""
    L = fd.lib()
    spec = [(0, 30), (1, 44), (2, 57)]

    def build():
        out = [_extruded(q, K) for q, K in spec]
        if repaint:
            fd.dev_check(L.fdtd_dev_paint_prism(out[1][0].d, 4.0, 1.0, 0, 4, 9, 6, 14, 12, 20), "paint")
        return out

    ref = []
    for g, mons, nf in build():
        g.run(GC.DT, 0, NSTEPS)
        ref.append(_results(g, mons, nf))
        g.free()
    built = build()
    b = fd.Batch([g for g, _, _ in built])
    b.set_shape(2, True, False)
    b.run(GC.DT, 0, 70)
    b.run(GC.DT, 70, NSTEPS - 70)
    for (g, mons, nf), (mon_s, f_s) in zip(built, ref):
        mon_b, f_b = _results(g, mons, nf)
        ef = float(np.linalg.norm(f_b - f_s) / np.linalg.norm(f_s))
        em = float(np.abs(mon_b - mon_s).max() / np.abs(mon_s).max())
        say("extruded (repaint %s): grid K %d fields rel %.2e, monitors max rel %.2e" % (repaint, g.K, ef, em))
        assert np.abs(mon_s).max() > 0 and ef <= 1e-6 and em <= 1e-6
    b.free()
    for g, _, _ in built:
        g.free()
