"""GPU tests of the extensions (no reference counterpart): CPML boundary and device-resident runs
with tabulated sources, FP64 overlap monitors and multi-frequency DFT planes.  The FP64 oracle
carries the same formulas (its CPML is pinned by the CPU tests in test_oracle_invariants.py);
gate: relative L2 <= 1e-4 for FP32 fields, same bound for FP64 monitors fed by FP32 fields."""
import ctypes as C
import importlib
import os
import sys

import numpy as np
import pytest

from oracle import oracle_py as O

pytestmark = pytest.mark.gpu

fd = importlib.import_module("3d_fdtd_amd")
L = fd.lib
REL = 1e-4
DP = C.POINTER(C.c_double)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_l2(got, want):
    return float(np.linalg.norm(np.asarray(got) - np.asarray(want)) / max(np.linalg.norm(want), 1e-300))


@pytest.mark.parametrize("npml", [(4, 3, 5, 2, 3, 4), (0, 0, 6, 6, 0, 5)])
def test_cpml_step_vs_oracle(npml):
    rng = np.random.default_rng(sum(npml))
    I, J, K, disc = 22, 18, 26, 0.02
    dt = 0.99 * disc / np.sqrt(3)
    say("cpml case", npml)
    g = fd.Grid(I, J, K, disc)
    o = O.Grid(I, J, K, disc)
    ep = rng.uniform(1, 12, 3 * I * J * K).astype(np.float32)
    sig = np.zeros(6 * I * J * K, np.float32)
    sig[0::6] = rng.uniform(0, 2, I * J * K).astype(np.float32)
    f0 = rng.uniform(-1, 1, 6 * I * J * K).astype(np.float32)
    for a, v in (("ep", ep), ("sig", sig), ("fields", f0)):
        g.a[a][:] = v
    o.ep[:] = ep
    o.sig[:] = sig
    o.f[:] = f0
    g.push()
    fd.check(L().fdtd3d_ext_set_cpml(g.tf, (C.c_int * 6)(*npml), 3, 1e-6, 2.0, 0.1), "set_cpml")
    cp = O.Cpml(o, npml, dt, order=3, R0=1e-6, kappa_max=2.0, alpha_max=0.1)
    for _ in range(40):
        L().timefield3d_apply_timestep_omp_full(g.tf, dt)
        O.step_cpml(o, cp, dt)
    err = rel_l2(g.fields(), o.f)
    say("cpml rel L2", err)
    assert err <= REL
    # zero_fields clears the PML memory too: a second identical run reproduces the first
    first = g.fields().copy()
    L().timefield3d_zero_fields(g.tf)
    g.a["fields"][:] = f0
    g.push()
    for _ in range(40):
        L().timefield3d_apply_timestep_omp_full(g.tf, dt)
    assert np.array_equal(g.fields(), first)
    g.free()


@pytest.mark.parametrize("rows,kchunk", [(-1, 1), (-1, 2), (-1, 4), (2, 3), (4, 16), (8, 5), (16, 3), (16, 7),
                                         (32, 4), (-3, 0), (-3, 1), (-3, 4), (-3, 7), (-3, 3)])
def test_fused_step_vs_split_and_oracle(rows, kchunk):
    """Fused one-kernel step (ping-pong buffers: columns, register strips, LDS tiles or the
    single-pass full step) vs the split E/H kernels and the FP64 oracle, on an odd-sized grid
    (partial blocks, strips, tiles and chunks) with conductivity, a plane source and PML on some
    faces; switches kernels mid-run to exercise the buffer resynchronisation."""
    rng = np.random.default_rng(abs(rows + 2) * 100 + kchunk)
    I, J, K, disc = 70, 23, 29, 0.02
    dt = 0.99 * disc / np.sqrt(3)
    npml = (5, 4, 0, 3, 6, 0)
    say("fused case rows", rows, "kchunk", kchunk)
    ep = rng.uniform(1, 12, 3 * I * J * K).astype(np.float32)
    sig = np.zeros(6 * I * J * K, np.float32)
    sig[1::6] = rng.uniform(0, 2, I * J * K).astype(np.float32)
    src = np.zeros(6 * I * J * K, np.float32)
    src.reshape(K, J, I, 6)[11] = rng.uniform(-1, 1, (J, I, 6)).astype(np.float32)
    f0 = rng.uniform(-1, 1, 6 * I * J * K).astype(np.float32)
    grids = []
    for mode in ("fused", "split", "mixed"):
        g = fd.Grid(I, J, K, disc)
        for a_, v in (("ep", ep), ("sig", sig), ("sources", src), ("fields", f0)):
            g.a[a_][:] = v
        g.push()
        fd.check(L().fdtd3d_ext_set_cpml(g.tf, (C.c_int * 6)(*npml), 3, 1e-6, 1.5, 0.1), "set_cpml")
        fd.check(L().fdtd3d_ext_set_step_kernel(g.tf, 0 if mode == "split" else rows, kchunk), "kernel")
        for s in range(30):
            if mode == "mixed" and s in (10, 20):
                fd.check(L().fdtd3d_ext_set_step_kernel(g.tf, 0 if s == 10 else rows, kchunk), "kernel")
            L().timefield3d_apply_timestep_omp_full(g.tf, dt)
        grids.append(g.fields().copy())
        g.free()
    o = O.Grid(I, J, K, disc)
    o.ep[:] = ep
    o.sig[:] = sig
    o.src[:] = src
    o.f[:] = f0
    cp = O.Cpml(o, npml, dt, order=3, R0=1e-6, kappa_max=1.5, alpha_max=0.1)
    for _ in range(30):
        O.step_cpml(o, cp, dt)
    for name, got in zip(("fused", "split", "mixed"), grids):
        err = rel_l2(got, o.f)
        say(name, "rel L2 vs oracle", err)
        assert err <= REL
    d = rel_l2(grids[0], grids[1])
    say("fused vs split", d)
    assert d <= 1e-6


@pytest.mark.parametrize("ring_monitor", [False, True])
def test_run_api_matches_reference_monitors(golden_dir, ring_monitor):
    """Device-resident run (tabulated source, overlap monitors with DFT) vs the reference's
    per-step source/overlap calls recorded in fixture F4 (single frequency).  The three F4
    monitors are accumulated inside the H update; an extra monitor whose window reaches the
    boundary ring switches the run to the separate monitor kernels."""
    say("run API vs F4, ring monitor", ring_monitor)
    z = np.load(os.path.join(golden_dir, "F4_mode_source_overlap.npz"))
    I, J, K = (int(v) for v in z["dims"])
    dt, om = float(z["dt"]), float(z["omega"])
    g = fd.Grid(I, J, K, float(z["disc"]))
    g.a["ep"][:] = z["ep"]
    g.a["mu"][:] = z["mu"]
    g.push()
    m1 = fd.mode_from_array(30, 30, float(z["disc"]), z["mode1"])
    m2 = fd.mode_from_array(24, 24, float(z["disc"]), z["mode2"])
    n = int(z["fwd_steps"])
    steps = np.arange(n)
    amp = np.exp(-(steps - 60) ** 2 / (2 * 25.0 ** 2))
    phase = steps * dt * om
    sid = L().fdtd3d_ext_run_add_mode_source(g.tf, I // 2, J // 2, 10, m1, 0, 0, n, amp.ctypes.data_as(DP),
                                             phase.ctypes.data_as(DP))
    assert sid >= 0
    oms = np.array([om])
    ids = [L().fdtd3d_ext_run_add_overlap(g.tf, I // 2, J // 2, kz, mm, rev, 1, oms.ctypes.data_as(DP))
           for kz, rev, mm in ((15, 0, m1), (15, 1, m1), (45, 0, m2))]
    assert min(ids) >= 0
    if ring_monitor:
        assert L().fdtd3d_ext_run_add_overlap(g.tf, 2, J // 2, 30, m1, 0, 1, oms.ctypes.data_as(DP)) >= 0
    # two calls, to exercise the step offset
    fd.check(L().fdtd3d_ext_run(g.tf, dt, 0, 50, 0.0), "run")
    fd.check(L().fdtd3d_ext_run(g.tf, dt, 50, n - 50, 0.0), "run")
    want = z["fwd_cum"]
    for q, mid in enumerate(ids):
        cum, last = fd.Complex(), fd.Complex()
        fd.check(L().fdtd3d_ext_run_overlap(g.tf, mid, C.byref(cum), C.byref(last)), "result")
        err = abs(complex(cum.re, cum.im) - complex(*want[q])) / np.abs(want).max()
        say("run monitor", q, complex(cum.re, cum.im), complex(*want[q]), err)
        assert err <= REL
        inst_err = abs(complex(last.re, last.im) - complex(*z["fwd_inst"][n - 1, q])) / np.abs(z["fwd_inst"][:, q]).max()
        assert inst_err <= REL
    # the caller's modes are untouched by the run API
    assert np.array_equal(fd.mode_fields(m1), z["mode1"])
    g.free()


def test_run_monitor_fold_matches_separate_kernels(golden_dir):
    """CPML grid, 16 frequencies, two run calls: monitors accumulated inside the H update agree
    with the separate monitor kernels (FP64 sums in another order: relative 1e-10)."""
    z = np.load(os.path.join(golden_dir, "F4_mode_source_overlap.npz"))
    I, J, K = (int(v) for v in z["dims"])
    dt, om = float(z["dt"]), float(z["omega"])
    n = 90
    steps = np.arange(n)
    amp = np.exp(-(steps - 40) ** 2 / (2 * 15.0 ** 2))
    phase = steps * dt * om
    oms = om * np.linspace(0.9, 1.1, 16)
    results = []
    for ring_monitor in (False, True):
        say("fold vs separate, ring monitor", ring_monitor)
        g = fd.Grid(I, J, K, float(z["disc"]))
        g.a["ep"][:] = z["ep"]
        g.a["mu"][:] = z["mu"]
        g.push()
        npml = (C.c_int * 6)(6, 6, 6, 6, 6, 6)
        fd.check(L().fdtd3d_ext_set_cpml(g.tf, npml, 3, 1e-6, 1.0, 0.05), "cpml")
        m1 = fd.mode_from_array(30, 30, float(z["disc"]), z["mode1"])
        m2 = fd.mode_from_array(24, 24, float(z["disc"]), z["mode2"])
        assert L().fdtd3d_ext_run_add_mode_source(g.tf, I // 2, J // 2, 10, m1, 0, 0, n, amp.ctypes.data_as(DP),
                                                  phase.ctypes.data_as(DP)) >= 0
        ids = [L().fdtd3d_ext_run_add_overlap(g.tf, I // 2, J // 2, kz, mm, rev, len(oms), oms.ctypes.data_as(DP))
               for kz, rev, mm in ((15, 0, m1), (15, 1, m1), (45, 0, m2))]
        assert min(ids) >= 0
        if ring_monitor:
            assert L().fdtd3d_ext_run_add_overlap(g.tf, 1, J // 2, 30, m1, 0, 2, oms.ctypes.data_as(DP)) >= 0
        fd.check(L().fdtd3d_ext_run(g.tf, dt, 0, 37, 0.0), "run")
        fd.check(L().fdtd3d_ext_run(g.tf, dt, 37, n - 37, 0.0), "run")
        out = []
        for mid in ids:
            cum = (fd.Complex * len(oms))()
            last = fd.Complex()
            fd.check(L().fdtd3d_ext_run_overlap(g.tf, mid, cum, C.byref(last)), "result")
            out.append([complex(c.re, c.im) for c in cum] + [complex(last.re, last.im)])
        results.append(np.array(out))
        g.free()
    fold, sep = results
    err = np.abs(fold - sep).max() / np.abs(sep).max()
    say("fold vs separate max rel", err)
    assert np.abs(sep).max() > 0
    assert err <= 1e-10


@pytest.mark.parametrize("rows", [-1, 8, 16, -3])
def test_run_api_fused_matches_split(golden_dir, rows):
    """Device-resident run under the PML with a fused step kernel (run sources evaluated inside
    the fused kernel; monitors inside the column kernel, by the separate kernels for strips and
    tiles) vs the split kernels (monitors inside the H update): overlaps agree to relative 1e-5."""
    z = np.load(os.path.join(golden_dir, "F4_mode_source_overlap.npz"))
    I, J, K = (int(v) for v in z["dims"])
    dt, om = float(z["dt"]), float(z["omega"])
    n = 120
    steps = np.arange(n)
    amp = np.exp(-(steps - 50) ** 2 / (2 * 20.0 ** 2))
    phase = steps * dt * om
    oms = om * np.linspace(0.9, 1.1, 8)
    results = []
    for kernel in (0, rows):
        say("run API step kernel rows", kernel)
        g = fd.Grid(I, J, K, float(z["disc"]))
        g.a["ep"][:] = z["ep"]
        g.a["mu"][:] = z["mu"]
        g.push()
        npml = (C.c_int * 6)(6, 6, 6, 6, 6, 6)
        fd.check(L().fdtd3d_ext_set_cpml(g.tf, npml, 3, 1e-6, 1.0, 0.05), "cpml")
        fd.check(L().fdtd3d_ext_set_step_kernel(g.tf, kernel, {0: 0, -1: 2, -3: 0}.get(kernel, 5)), "kernel")
        m1 = fd.mode_from_array(30, 30, float(z["disc"]), z["mode1"])
        assert L().fdtd3d_ext_run_add_mode_source(g.tf, I // 2, J // 2, 10, m1, 0, 0, n, amp.ctypes.data_as(DP),
                                                  phase.ctypes.data_as(DP)) >= 0
        ids = [L().fdtd3d_ext_run_add_overlap(g.tf, I // 2, J // 2, kz, m1, rev, len(oms), oms.ctypes.data_as(DP))
               for kz, rev in ((15, 0), (15, 1), (45, 0))]
        fd.check(L().fdtd3d_ext_run(g.tf, dt, 0, n, 0.0), "run")
        out = []
        for mid in ids:
            cum = (fd.Complex * len(oms))()
            fd.check(L().fdtd3d_ext_run_overlap(g.tf, mid, cum, None), "result")
            out.append([complex(c.re, c.im) for c in cum])
        results.append(np.array(out))
        g.free()
    err = np.abs(results[0] - results[1]).max() / np.abs(results[0]).max()
    say("fused vs split run monitors max rel", err)
    assert np.abs(results[0]).max() > 0
    assert err <= 1e-5


def test_run_long_call_matches_split_calls(golden_dir):
    """A run call longer than the library's internal step chunk (4096) agrees with the same
    steps issued as two calls with other chunk boundaries (monitors to relative 1e-10)."""
    z = np.load(os.path.join(golden_dir, "F4_mode_source_overlap.npz"))
    I, J, K = (int(v) for v in z["dims"])
    dt, om = float(z["dt"]), float(z["omega"])
    n = 4200
    steps = np.arange(n)
    amp = np.exp(-(steps - 60) ** 2 / (2 * 25.0 ** 2))
    phase = steps * dt * om
    oms = om * np.linspace(0.95, 1.05, 4)
    results = []
    for calls in ((0, n), (0, 100, n)):
        say("long run, call boundaries", calls)
        g = fd.Grid(I, J, K, float(z["disc"]))
        g.a["ep"][:] = z["ep"]
        g.a["mu"][:] = z["mu"]
        g.push()
        npml = (C.c_int * 6)(6, 6, 6, 6, 6, 6)
        fd.check(L().fdtd3d_ext_set_cpml(g.tf, npml, 3, 1e-6, 1.0, 0.05), "cpml")
        m1 = fd.mode_from_array(30, 30, float(z["disc"]), z["mode1"])
        assert L().fdtd3d_ext_run_add_mode_source(g.tf, I // 2, J // 2, 10, m1, 0, 0, n, amp.ctypes.data_as(DP),
                                                  phase.ctypes.data_as(DP)) >= 0
        mid = L().fdtd3d_ext_run_add_overlap(g.tf, I // 2, J // 2, 45, m1, 0, len(oms), oms.ctypes.data_as(DP))
        assert mid >= 0
        for a, b in zip(calls[:-1], calls[1:]):
            fd.check(L().fdtd3d_ext_run(g.tf, dt, a, b - a, 0.0), "run")
        cum = (fd.Complex * len(oms))()
        last = fd.Complex()
        fd.check(L().fdtd3d_ext_run_overlap(g.tf, mid, cum, C.byref(last)), "result")
        results.append(np.array([complex(c.re, c.im) for c in cum] + [complex(last.re, last.im)]))
        g.free()
    err = np.abs(results[0] - results[1]).max() / np.abs(results[1]).max()
    say("long vs split calls max rel", err)
    assert np.abs(results[1]).max() > 0
    assert err <= 1e-10


def test_run_dft_plane_vs_oracle(golden_dir):
    """Multi-frequency FP64 DFT plane vs the oracle's reference-semantics DFT on a plane view."""
    z = np.load(os.path.join(golden_dir, "F4_mode_source_overlap.npz"))
    I, J, K = (int(v) for v in z["dims"])
    dt, om = float(z["dt"]), float(z["omega"])
    n = 80
    g = fd.Grid(I, J, K, float(z["disc"]))
    g.a["ep"][:] = z["ep"]
    g.a["mu"][:] = z["mu"]
    g.push()
    m1 = fd.mode_from_array(30, 30, float(z["disc"]), z["mode1"])
    steps = np.arange(n)
    amp = np.exp(-(steps - 40) ** 2 / (2 * 15.0 ** 2))
    phase = steps * dt * om
    assert L().fdtd3d_ext_run_add_mode_source(g.tf, I // 2, J // 2, 10, m1, 0, 0, n, amp.ctypes.data_as(DP),
                                              phase.ctypes.data_as(DP)) >= 0
    oms = np.array([om, 0.8 * om])
    pid = L().fdtd3d_ext_run_add_dft_plane(g.tf, 30, 0b010011, len(oms), oms.ctypes.data_as(DP))
    assert pid >= 0
    fd.check(L().fdtd3d_ext_run(g.tf, dt, 0, n, 0.0), "run")
    out = (fd.Complex * (3 * len(oms) * I * J))()
    fd.check(L().fdtd3d_ext_run_dft(g.tf, pid, out), "dft result")
    got = np.ctypeslib.as_array(C.cast(out, DP), shape=(3, len(oms), J, I, 2))
    # oracle
    o = O.Grid(I, J, K, float(z["disc"]))
    o.ep[:] = z["ep"]
    o.mu[:] = z["mu"]
    mo = z["mode1"].copy()
    P = I * J
    bufs = [np.zeros(12 * P) for _ in oms]
    for s in range(n):
        O.lib().orc_mode_source(o.ptr, I // 2, J // 2, 10, 30, 30, o.disc, O.dptr(mo), phase[s], amp[s], 0)
        o.step(dt)
        view_f = o.f[6 * P * 30: 6 * P * 31]
        vs = O.OrcGrid(I, J, 1, o.disc, O.fptr(o.ep), O.fptr(o.mu), O.fptr(o.sig), O.dptr(view_f), O.dptr(o.old),
                       O.fptr(o.src))
        for q, w in enumerate(oms):
            O.lib().orc_dft_volume(C.byref(vs), s * dt * w, O.dptr(bufs[q]))
    want = np.stack([b.reshape(J, I, 6, 2) for b in bufs])  # [f, j, i, m, 2]
    want = np.stack([want[:, :, :, m, :] for m in (0, 1, 4)])  # [comp, f, j, i, 2]
    err = rel_l2(got, want)
    say("run DFT plane rel L2", err)
    assert err <= REL
    g.free()


def test_nonfinite_fields_are_detected():
    """A NaN planted in the fields from the host (failure injection; every kernel still runs to
    completion on NaN data) is counted by the one-shot scan and fails a watched run with -4,
    naming the first scanned step that showed it; clean fields pass both."""
    say("non-finite watch")
    I, J, K, disc = 40, 30, 36, 0.02
    dt = 0.99 * disc / np.sqrt(3)
    ep = np.full(3 * I * J * K, 2.25, np.float32)
    g = fd.DevGrid(I, J, K, disc, ep)
    g.cpml((4,) * 6, 3, 1e-8, 1.0, 0.05)
    f0 = np.random.default_rng(3).uniform(-1, 1, 6 * I * J * K)
    g.set_fields(f0)
    n = C.c_longlong(-1)
    fd.dev_check(L().fdtd_dev_count_nonfinite(g.d, C.byref(n)), "count")
    assert n.value == 0
    fd.dev_check(L().fdtd_dev_set_finite_check(g.d, 10), "watch")
    assert L().fdtd_dev_run(g.d, dt, 0, 30, 0.0) == 0
    bad = g.fields()
    c = 6 * ((K // 2 * J + J // 2) * I + I // 2) + 2  # Ez at the centre
    bad[c] = np.nan
    g.set_fields(bad)
    fd.dev_check(L().fdtd_dev_count_nonfinite(g.d, C.byref(n)), "count")
    assert n.value == 1
    rc = L().fdtd_dev_run(g.d, dt, 30, 30, 0.0)
    msg = L().fdtd_dev_last_error().decode()
    say("watched run:", rc, msg)
    assert rc == -4 and "after step 39" in msg
    fd.dev_check(L().fdtd_dev_count_nonfinite(g.d, C.byref(n)), "count")
    assert n.value > 100  # the NaN has spread through the stencil
    g.free()


def test_streaming_loads_change_nothing_but_speed():
    """The split updates with and without the streaming hint on the own-cell loads (forced on and
    off on a small grid): identical fields after a CPML run with a plane source."""
    say("streaming loads on / off")
    I, J, K, disc = 34, 26, 30, 0.02
    dt = 0.99 * disc / np.sqrt(3)
    rng = np.random.default_rng(8)
    ep = rng.uniform(1, 12, 3 * I * J * K).astype(np.float32)
    f0 = rng.uniform(-1, 1, 6 * I * J * K)
    out = []
    for mode in (0, 1):
        g = fd.DevGrid(I, J, K, disc, ep)
        g.cpml((4, 3, 5, 2, 3, 4), 3, 1e-6, 1.5, 0.1)
        fd.dev_check(L().fdtd_dev_set_streaming_loads(g.d, mode), "streaming")
        g.set_fields(f0)
        g.run(dt, 0, 25)
        out.append(g.fields())
        g.free()
    assert np.abs(out[0]).max() > 0
    assert np.array_equal(out[0], out[1])
    g = fd.DevGrid(I, J, K, disc, ep)
    assert L().fdtd_dev_set_streaming_loads(g.d, 2) == -2
    g.free()
