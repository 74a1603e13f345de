"""GPU parity of the drop-in path (FP32 HIP kernels) against the reference (FP64).

Every test drives libfdtd3d.so through its C-ABI (the reference's own call surface) and compares
with the committed fixtures written by the compiled reference, or with the FP64 oracle at sizes
it finishes in seconds.  Gates: relative L2 field error <= 1e-4 (FP32 fields vs FP64 reference);
FP64 monitor/gradient paths fed from FP32 fields are held to the same 1e-4 relative bound, and
reductions over exactly representable FP32 data to a few ulp.
"""
import ctypes as C
import hashlib
import importlib
import os
import sys

import numpy as np
import pytest

from oracle import oracle_py as O

pytestmark = pytest.mark.gpu

fd = importlib.import_module("3d_fdtd_amd")
L = fd.lib
REL_L2 = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_l2(got, want):
    return float(np.linalg.norm(got - want) / max(np.linalg.norm(want), 1e-300))


def load(golden_dir, name):
    return np.load(os.path.join(golden_dir, name))


def grid_from(z, I, J, K):
    g = fd.Grid(I, J, K, float(z["disc"]))
    g.a["ep"][:] = z["ep"]
    g.a["mu"][:] = z["mu"]
    return g


@pytest.mark.parametrize("name", ["F1_step_random.npz", "F2_step_sigma_sources.npz"])
def test_step_random_vs_reference(golden_dir, name):
    z = load(golden_dir, name)
    I, J, K = (int(v) for v in z["dims"])
    say("case", name, I, J, K)
    g = grid_from(z, I, J, K)
    g.a["fields"][:] = z["fields0"]
    if "sig" in z:
        g.a["sig"][:] = z["sig"]
        g.a["sources"][:] = z["sources"]
    g.push()
    g.step(float(z["dt"]), int(z["steps"]))
    got = g.fields().copy()
    err = rel_l2(got, z["fields"])
    say(name, "rel L2", err)
    assert err <= REL_L2
    g.free()


def test_zsection_entry_points_match_oracle():
    """The _zsection E/H updates and Mur faces driven slab by slab, vs the oracle's full step."""
    rng = np.random.default_rng(7)
    I, J, K, disc = 18, 14, 20, 0.02
    dt = 0.99 * disc / np.sqrt(3)
    g = fd.Grid(I, J, K, disc)
    o = O.Grid(I, J, K, disc)
    for key, lo, hi in (("ep", 1, 12), ("mu", 1, 2)):
        v = rng.uniform(lo, hi, g.a[key].size).astype(np.float32)
        g.a[key][:] = v
        getattr(o, key)[:] = v
    f0 = rng.uniform(-1, 1, g.a["fields"].size).astype(np.float32)
    g.a["fields"][:] = f0
    o.f[:] = f0
    g.push()
    # reference-mode step assembled from the zsection entry points over two slabs, with the
    # one-plane widening of the reference's OpenMP driver (timefield3d.c:2794-2808)
    for _ in range(10):
        L().timefield3d_update_old_edge_fields(g.tf)
        for lo, hi in ((0, 11), (9, K)):
            L().timefield3d_apply_E_timestep_zsection(g.tf, dt, lo, hi)
        L().timefield3d_apply_E_first_order_abc_zsection(g.tf, dt, 0, 10)
        L().timefield3d_apply_E_first_order_abc_zsection(g.tf, dt, 10, K)
        for lo, hi in ((0, 11), (9, K)):
            L().timefield3d_apply_H_timestep_zsection(g.tf, dt, lo, hi)
        L().timefield3d_apply_H_first_order_abc_zsection(g.tf, dt, 0, 10)
        L().timefield3d_apply_H_first_order_abc_zsection(g.tf, dt, 10, K)
        o.step(dt)
    err = rel_l2(g.fields(), o.f)
    say("zsection assembly rel L2", err)
    assert err <= REL_L2
    g.free()


def test_explicit_step_vs_reference(golden_dir):
    """F7: timefield3d_apply_timestep (explicit conductivity, timefield3d.c:1537-1677)."""
    z = load(golden_dir, "F7_explicit_block.npz")
    I, J, K = (int(v) for v in z["dims"])
    g = grid_from(z, I, J, K)
    g.a["sig"][:] = z["sig"]
    g.a["sources"][:] = z["sources"]
    g.a["fields"][:] = z["fields0"]
    g.push()
    for _ in range(int(z["steps"])):
        L().timefield3d_apply_timestep(g.tf, float(z["dt"]))
    err = rel_l2(g.fields(), z["fields"])
    say("explicit step rel L2", err)
    assert err <= REL_L2
    g.free()


def test_block_adjoint_vs_reference(golden_dir):
    """F7: 3-D block adjoint store / reverse accumulation / permittivity step / taper projection.
    The stored values are exact FP32 numbers, so the FP64 device accumulation is bit-exact."""
    z = load(golden_dir, "F7_explicit_block.npz")
    BI, BJ, BK = (int(v) for v in z["block_dims"])
    box, T = [int(v) for v in z["box"]], int(z["T"])
    g = fd.Grid(BI, BJ, BK, 0.02)
    blk = L().ajoint3d_new(*box, T)
    fwd = pattern_fields(T, g.a["fields"].size, 7, 13)
    adj = pattern_fields(T, g.a["fields"].size, 11, 5)
    for t in range(T):
        g.a["fields"][:] = fwd[t]
        g.push()
        L().adjoint3d_store_fwd_fields(blk, g.tf, t)
    for t in range(T - 1, -1, -1):
        g.a["fields"][:] = adj[t]
        g.push()
        L().adjoint3d_update_gradient(blk, g.tf, t)
    nb = (box[1] - box[0]) * (box[3] - box[2]) * (box[5] - box[4])
    grad = np.ctypeslib.as_array(blk.gradient, shape=(nb,))
    assert np.array_equal(grad, z["grad"])
    L().timefield3d_set_background(g.tf, 2.0, 1.0)
    L().adjoint3d_adjust_permittivity(blk, g.tf, 4.0, 0.05, 1.5, 12.0)
    assert np.array_equal(g.a["ep"], z["ep_after"])
    left = np.ascontiguousarray(z["taper_left"].astype(np.int32))
    right = np.ascontiguousarray(z["taper_right"].astype(np.int32))
    ip = C.POINTER(C.c_int)
    tp = L().adjoint3d_taper_new(12.0, 1.0, 3, 5, 14, left.ctypes.data_as(ip), right.ctypes.data_as(ip), 3, 9, 1)
    L().adjoint3d_taper_calculate_gradients(tp, blk, 10.0)
    assert np.array_equal(np.ctypeslib.as_array(tp.left_grad, shape=(4,)), z["taper_left_grad"])
    assert np.array_equal(np.ctypeslib.as_array(tp.right_grad, shape=(4,)), z["taper_right_grad"])
    g.free()


def dipole_J(n, dt):
    lam0, dlam = 1.55, 0.2
    w0 = 2 * np.pi / lam0
    sig_t = lam0 * lam0 / (2 * np.pi * dlam)
    t = n * dt
    return np.exp(-(t - 250 * dt) ** 2 / (2 * sig_t ** 2)) * np.sin(w0 * t)


def test_dipole_mur_64(golden_dir):
    """64^3 vacuum point dipole, Mur, 500 steps: Ez planes vs the reference fixture, full fields
    vs the FP64 oracle, and the plane-48 Ez DFT at 5 wavelengths vs the fixture."""
    z = load(golden_dir, "F3_dipole_mur_64.npz")
    N = 64
    dt = float(z["dt"])
    c0 = 32 + N * (32 + N * 32)
    dfts = []
    for q, w in enumerate(z["omegas"]):
        g = fd.Grid(N, N, N, float(z["disc"]))
        for n in range(int(z["steps"])):
            g.a["sources"][6 * c0 + 2] = z["J"][n]
            fd.check(L().fdtd3d_ext_upload_source_plane(g.tf, 32), "source plane")
            g.step(dt)
            L().timefield3d_update_fourier_fields_zsection(g.tf, n * dt * float(w), 48, 49)
        g.pull()
        F = g.a["fields"].reshape(N, N, N, 6)
        if q == 0:
            say("dipole Ez z=32 rel L2", rel_l2(F[32, :, :, 2], z["ez_z32"]))
            assert rel_l2(F[32, :, :, 2], z["ez_z32"]) <= REL_L2
            assert rel_l2(F[48, :, :, 2], z["ez_z48"]) <= REL_L2
            full = F.copy()
        four = g.a["fourier"].reshape(N, N, N, 6, 2)[48, :, :, 2, :]
        dfts.append(four.copy())
        g.free()
    err = rel_l2(np.stack(dfts), z["dft_ez_z48"])
    say("dipole DFT rel L2", err)
    assert err <= REL_L2
    # full six-component field vs the FP64 oracle (bit-exact to the reference on this fixture)
    o = O.Grid(N, N, N, float(z["disc"]))
    for n in range(int(z["steps"])):
        o.src[6 * c0 + 2] = z["J"][n]
        o.step(dt)
    err = rel_l2(full.reshape(-1), o.f)
    say("dipole full-field rel L2 vs oracle", err)
    assert err <= REL_L2


def test_dipole_cpml_64():
    """The CPU-runnable configuration as specified, with its 8-cell CPML on all faces: 64^3
    vacuum, 20 nm cells, z-polarised point dipole at the centre driven by the Gaussian-modulated
    sine (lambda0 1.55 um, bandwidth 0.2 um), Courant 0.99/sqrt(3), 500 steps.  Full Ez at the
    last step and the running DFT of Ez on the plane z = 48 at 5 wavelengths 1.50-1.60 um (the
    drop-in per-frequency accumulator, one run per wavelength) against the FP64 oracle."""
    N, disc, steps = 64, 0.02, 500
    dt = 0.99 * disc / np.sqrt(3)
    npml = (8,) * 6
    oms = 2 * np.pi / np.linspace(1.50, 1.60, 5)
    J = dipole_J(np.arange(steps), dt)
    c0 = 32 + N * (32 + N * 32)
    say("64^3 dipole under the CPML, 500 steps x 5 wavelengths")
    dft_gpu, ez_gpu = [], None
    for q, w in enumerate(oms):
        g = fd.Grid(N, N, N, disc)
        fd.check(L().fdtd3d_ext_set_cpml(g.tf, (C.c_int * 6)(*npml), 3, 1e-8, 1.0, 0.05), "cpml")
        for n in range(steps):
            g.a["sources"][6 * c0 + 2] = J[n]
            fd.check(L().fdtd3d_ext_upload_source_plane(g.tf, 32), "source plane")
            g.step(dt)
            L().timefield3d_update_fourier_fields_zsection(g.tf, n * dt * float(w), 48, 49)
        g.pull()
        if q == 0:
            ez_gpu = g.a["fields"].reshape(N, N, N, 6)[..., 2].copy()
        f = g.a["fourier"].reshape(N, N, N, 6, 2)[48, :, :, 2, :]
        dft_gpu.append(f[..., 0] + 1j * f[..., 1])
        g.free()
    o = O.Grid(N, N, N, disc)
    cp = O.Cpml(o, npml, dt, order=3, R0=1e-8, kappa_max=1.0, alpha_max=0.05)
    dft_ref = np.zeros((len(oms), N, N), complex)
    for n in range(steps):
        o.src[6 * c0 + 2] = J[n]
        O.step_cpml(o, cp, dt)
        dft_ref += np.exp(1j * oms * n * dt)[:, None, None] * o.comp(2)[48][None]
    e_ez = rel_l2(ez_gpu, o.comp(2))
    e_dft = rel_l2(np.stack(dft_gpu), dft_ref)
    say("CPML dipole: Ez rel L2", e_ez, "DFT(z=48, 5 wavelengths) rel L2", e_dft)
    assert np.abs(o.comp(2)).max() > 0 and np.abs(dft_ref).max() > 0
    assert e_ez <= REL_L2
    assert e_dft <= REL_L2


def test_mode_source_overlap_adjoint(golden_dir):
    """F4: mode source + y-slice DFT + three overlap monitors, then the CW adjoint run."""
    z = load(golden_dir, "F4_mode_source_overlap.npz")
    I, J, K = (int(v) for v in z["dims"])
    dt, om = float(z["dt"]), float(z["omega"])
    g = grid_from(z, I, J, K)
    m1 = fd.mode_from_array(30, 30, float(z["disc"]), z["mode1"])
    m2 = fd.mode_from_array(24, 24, float(z["disc"]), z["mode2"])
    mons = [(15, 0, m1), (15, 1, m1), (45, 0, m2)]
    cums = [fd.Complex(0, 0) for _ in mons]
    inst = np.zeros((int(z["fwd_steps"]), 3, 2))
    for n in range(int(z["fwd_steps"])):
        amp = np.exp(-(n - 60) ** 2 / (2 * 25.0 ** 2))
        ph = n * dt * om
        L().timefield3d_mode_source(g.tf, I // 2, J // 2, 10, m1, ph, amp, 0)
        L().timefield3d_apply_timestep_omp_full(g.tf, dt)
        L().timefield3d_update_fourier_fields_y_slice_omp(g.tf, ph, J // 2)
        for q, (kz, rev, mm) in enumerate(mons):
            r = L().timefield3d_mode_overlap(g.tf, I // 2, J // 2, kz, mm, ph, rev, C.byref(cums[q]))
            inst[n, q] = (r.re, r.im)
    want = z["fwd_inst"]
    for q in range(3):
        e = rel_l2(inst[:, q], want[:, q])
        say("overlap monitor", q, "rel L2 over steps", e)
        assert e <= REL_L2
    cum = np.array([[c.re, c.im] for c in cums])
    say("cumulative", cum, z["fwd_cum"])
    assert np.all(np.abs(cum - z["fwd_cum"]) <= REL_L2 * np.abs(z["fwd_cum"]).max())
    # the caller's mode is renormalised in place exactly like the reference
    assert np.array_equal(fd.mode_fields(m1), z["mode1_after_fwd"])
    a = g.pull()
    assert np.array_equal(a["sources"].reshape(K, J, I, 6)[10], z["fwd_src_plane"])
    ex_plane = a["fields"].reshape(K, J, I, 6)[:, J // 2, :, 0]
    assert rel_l2(ex_plane, z["fwd_ex_plane"]) <= REL_L2
    dft = a["fourier"].reshape(K, J, I, 6, 2)[:, J // 2, :, 0:2, :]
    say("y-slice DFT rel L2", rel_l2(dft, z["fwd_dft_y_exey"]))
    assert rel_l2(dft, z["fwd_dft_y_exey"]) <= REL_L2
    # adjoint
    L().mode_smul(fd.Complex(0.3, 0.8), m1)
    assert np.array_equal(fd.mode_fields(m1), z["mode1_rot"])
    L().timefield3d_zero_fields(g.tf)
    for n in range(int(z["adj_steps"]), 0, -1):
        L().timefield3d_mode_source_adjoint(g.tf, I // 2, J // 2, 45, m1, (n + 1) * dt * om, 0)
        L().timefield3d_apply_timestep_omp_full(g.tf, dt)
    a = g.pull()
    assert np.array_equal(a["sources"].reshape(K, J, I, 6)[45], z["adj_src_plane"])
    hy = a["fields"].reshape(K, J, I, 6)[:, J // 2, :, 4]
    say("adjoint Hy plane rel L2", rel_l2(hy, z["adj_hy_plane"]))
    assert rel_l2(hy, z["adj_hy_plane"]) <= REL_L2
    g.free()


def pattern_fields(T, n, a, b):
    idx = np.arange(n, dtype=np.int64)
    return np.stack([(((a * idx + b * t) % 97) / 48.5 - 1.0) for t in range(T)]).astype(np.float32)


def test_taper_store_and_gradients(golden_dir):
    """F6: device edge histories and the FP64 device correlation vs the reference; the stored
    values are exact FP32 numbers, so the gradient matches to a few ulp."""
    z = load(golden_dir, "F6_taper.npz")
    I, J, K = (int(v) for v in z["dims"])
    seg, start, end, bottom, top, T = (int(z[k]) for k in ("seg", "start", "end", "bottom", "top", "T"))
    left = np.ascontiguousarray(z["left"].astype(np.int32))
    right = np.ascontiguousarray(z["right"].astype(np.int32))
    ip = C.POINTER(C.c_int)
    tp = L().adjoint3d_taper_new(3.51 ** 2, 1.0, seg, start, end, left.ctypes.data_as(ip), right.ctypes.data_as(ip),
                                 bottom, top, T)
    g = fd.Grid(I, J, K, 0.02)
    for store, (a, b) in ((L().adjoint3d_taper_store_forward_fields, (7, 13)),
                          (L().adjoint3d_taper_store_adjoint_fields, (11, 5))):
        seq = pattern_fields(T, g.a["fields"].size, a, b)
        for t in range(T):
            g.a["fields"][:] = seq[t]
            g.push()
            store(tp, g.tf, t)
    L().adjoint3d_taper_calculate_ep_gradient(tp)
    H, Ln = top - bottom, end - start
    epg = np.ctypeslib.as_array(tp.ep_gradient, shape=(2 * H * Ln,)).copy()
    want = z["ep_grad"]
    err = np.abs(epg - want).max() / np.abs(want).max()
    say("ep gradient max rel", err)
    assert err <= 1e-13
    L().adjoint3d_taper_calculate_gradients_from_ep_gradient(tp, float(z["ep_diff"]))
    lg = np.ctypeslib.as_array(tp.left_grad, shape=(seg + 1,))
    rg = np.ctypeslib.as_array(tp.right_grad, shape=(seg + 1,))
    assert np.allclose(lg, z["left_grad"], rtol=1e-12, atol=0) and np.allclose(rg, z["right_grad"], rtol=1e-12, atol=0)
    g.free()


def test_block_adjust_over_65536_cells_vs_oracle():
    """A design step of the block adjoint over 76,000 cells gives nearly every cell of the block its own
    permittivity (random forward/adjoint fields, so every gradient value differs): more distinct
    (ep, mu, sigma) tuples than 16-bit material ids hold.  The drop-in step then runs on 4-byte ids
    and must match the FP64 oracle on the same permittivity over 30 steps."""
    I, J, K = 56, 46, 44
    box = (3, 53, 3, 43, 3, 41)  # 50 x 40 x 38 = 76,000 cells
    T = 2
    rng = np.random.default_rng(6553)
    g = fd.Grid(I, J, K, 0.02)
    blk = L().ajoint3d_new(*box, T)
    n6 = g.a["fields"].size
    for t in range(T):
        g.a["fields"][:] = rng.uniform(-1, 1, n6)
        g.push()
        L().adjoint3d_store_fwd_fields(blk, g.tf, t)
    for t in range(T - 1, -1, -1):
        g.a["fields"][:] = rng.uniform(-1, 1, n6)
        g.push()
        L().adjoint3d_update_gradient(blk, g.tf, t)
    L().timefield3d_set_background(g.tf, 2.0, 1.0)
    # the step moves the permittivity by improvement * g / (|g|^2 disc^2 omega): 0.1 rms here, so
    # nearly every cell takes its own float value (a smaller step repeats values at the float
    # spacing near 2, a larger one clips at the bounds)
    L().adjoint3d_adjust_permittivity(blk, g.tf, 4.0, 10.0, 1.5, 12.0)
    ep = g.a["ep"].copy()
    distinct = len(np.unique(ep.reshape(-1, 3), axis=0))
    say("distinct permittivity tuples after the design step:", distinct)
    assert distinct > 65536
    f0 = rng.uniform(-1, 1, n6)
    g.a["fields"][:] = f0
    g.push()
    dt = 0.99 * 0.02 / np.sqrt(3)
    g.step(dt, 30)
    got = g.fields().copy()
    o = O.Grid(I, J, K, 0.02)
    o.ep[:] = ep
    o.f[:] = f0
    for _ in range(30):
        o.step(dt)
    err = rel_l2(got, o.f)
    say("graded block, 30 drop-in steps: rel L2 vs FP64 oracle", err)
    assert err <= REL_L2
    L().adjoint3d_free(blk)
    g.free()


def test_gaussian_sources_vs_reference(golden_dir):
    """F13: TE and TM Gaussian sources written through the drop-in surface equal the reference's
    whole source array bit for bit (same double arithmetic, rounded once to float)."""
    z = load(golden_dir, "F13_gaussian_sources.npz")
    I, J, K = (int(v) for v in z["dims"])
    g = fd.Grid(I, J, K, float(z["disc"]))
    g.a["ep"][:] = z["ep"]
    g.a["mu"][:] = z["mu"]
    g.a["sig"][:] = z["sig"]
    om = float(z["omega"])
    te, tm = z["te"], z["tm"]
    L().timefield3d_TE_gaussian_source(g.tf, int(te[0]), int(te[1]), int(te[2]), float(te[3]), om, float(te[4]),
                                       float(te[5]))
    L().timefield3d_TM_gaussian_source(g.tf, int(tm[0]), int(tm[1]), int(tm[2]), float(tm[3]), om, float(tm[4]),
                                       float(tm[5]))
    got = g.pull()["sources"].copy()
    diff = int(np.count_nonzero(got != z["sources"]))
    say("gaussian sources: values differing from the reference:", diff, "of", got.size)
    assert np.array_equal(got, z["sources"])
    g.free()


def test_device_raster_of_reference_setters_vs_reference(golden_dir):
    """F6's raster (background, top half plane, set_taper, adjoint3d_taper_insert, set_waveguide)
    with the last three applied to the device ids by repaints (the device materials being current
    after the first two): the device's per-cell ep and mu equal the reference's raster bit for bit,
    with one full upload in total, and steps on the repainted grid equal steps on a grid whose
    materials were uploaded from the same host arrays."""
    z = load(golden_dir, "F6_taper.npz")
    seg, start, end, bottom, top = (int(z[k]) for k in ("seg", "start", "end", "bottom", "top"))
    RI, RJ, RK = (int(v) for v in z["raster_dims"])
    left = np.ascontiguousarray(z["left"].astype(np.int32))
    right = np.ascontiguousarray(z["right"].astype(np.int32))
    ip = C.POINTER(C.c_int)
    counts = (C.c_longlong * 3)()
    g = fd.Grid(RI, RJ, RK, 0.02)
    L().timefield3d_set_background(g.tf, 1.447 ** 2, 1.0)
    L().timefield3d_set_top_half_plane(g.tf, 1.93 ** 2, 1.0, 20)
    f0 = np.random.default_rng(66).uniform(-1, 1, g.a["fields"].size)
    g.a["fields"][:] = f0
    g.push()  # fields and materials to the device: its materials are current from here on
    fd.check(L().fdtd3d_ext_material_counts(g.tf, counts), "material counts")
    assert list(counts)[:2] == [1, 0]
    L().timefield3d_set_taper(g.tf, 3.51 ** 2, 1.0, 20, 20, 14, 26, 0, 10, 12, 16)
    tp = L().adjoint3d_taper_new(3.51 ** 2, 1.0, seg, start, end, left.ctypes.data_as(ip), right.ctypes.data_as(ip),
                                 bottom, top, 1)
    L().adjoint3d_taper_insert(g.tf, tp)
    L().timefield3d_set_waveguide(g.tf, 3.51 ** 2, 10, 30, 35, 50, 12, 16)
    fd.check(L().fdtd3d_ext_material_counts(g.tf, counts), "material counts")
    say("material uploads, device repaints, table entries:", list(counts))
    assert counts[0] == 1 and counts[1] == 1 + seg + 1
    n = RI * RJ * RK
    ep = np.zeros(3 * n, np.float32)
    mu = np.zeros(3 * n, np.float32)
    fp = C.POINTER(C.c_float)
    fd.check(L().fdtd3d_ext_device_materials(g.tf, ep.ctypes.data_as(fp), mu.ctypes.data_as(fp), None), "materials")
    assert np.array_equal(ep, z["raster_ep"]) and np.array_equal(mu, z["raster_mu"])
    assert np.array_equal(g.a["ep"], z["raster_ep"]) and np.array_equal(g.a["mu"], z["raster_mu"])
    # stepping on the repainted ids equals stepping on the same permittivity uploaded from the host
    h = fd.Grid(RI, RJ, RK, 0.02)
    h.a["ep"][:] = z["raster_ep"]
    h.a["mu"][:] = z["raster_mu"]
    h.a["fields"][:] = f0
    h.push()
    dt = 0.99 * 0.02 / np.sqrt(3)
    g.step(dt, 20)
    h.step(dt, 20)
    fd.check(L().fdtd3d_ext_material_counts(g.tf, counts), "material counts")
    assert counts[0] == 1  # the steps used the repainted ids, no upload
    assert np.array_equal(g.fields(), h.fields())
    g.free()
    h.free()


@pytest.mark.parametrize("dims", [(18, 14, 20), (5, 3, 4), (33, 17, 9)], ids=["18x14x20", "5x3x4", "33x17x9"])
def test_one_launch_mur_equals_face_by_face(dims):
    """The full step applies all six Mur faces of a field in one launch (k_mur_all), resolving the
    edges and corners by the reference's face order inside the kernel; the zsection assembly runs
    the faces one launch each, in that order.  Random materials and fields, 10 steps: the two must
    be bit-identical, including on grids only 3 cells thick."""
    I, J, K = dims
    rng = np.random.default_rng(1 + I)
    dt = 0.99 * 0.02 / np.sqrt(3)
    grids = [fd.Grid(I, J, K, 0.02) for _ in range(2)]
    ep = rng.uniform(1, 12, grids[0].a["ep"].size).astype(np.float32)
    mu = rng.uniform(1, 2, grids[0].a["mu"].size).astype(np.float32)
    f0 = rng.uniform(-1, 1, grids[0].a["fields"].size)
    for g in grids:
        g.a["ep"][:] = ep
        g.a["mu"][:] = mu
        g.a["fields"][:] = f0
        g.push()
    a, b = grids
    h = K // 2
    for _ in range(10):
        L().timefield3d_apply_timestep_omp_full(a.tf, dt)
        L().timefield3d_update_old_edge_fields(b.tf)
        L().timefield3d_apply_E_timestep_zsection(b.tf, dt, 0, K)
        L().timefield3d_apply_E_first_order_abc_zsection(b.tf, dt, 0, h)
        L().timefield3d_apply_E_first_order_abc_zsection(b.tf, dt, h, K)
        L().timefield3d_apply_H_timestep_zsection(b.tf, dt, 0, K)
        L().timefield3d_apply_H_first_order_abc_zsection(b.tf, dt, 0, h)
        L().timefield3d_apply_H_first_order_abc_zsection(b.tf, dt, h, K)
    fa, fb = a.fields().copy(), b.fields().copy()
    say(f"{dims}: values differing between one-launch and face-by-face Mur:", int(np.count_nonzero(fa != fb)))
    assert np.array_equal(fa, fb)
    for g in grids:
        g.free()
