"""Pin the clean-room FP64 oracle to the reference: bit-exact against the committed fixtures
(written by tests/golden/make_golden.py from the compiled reference), plus invariants for the
parts no fixture covers.  CPU only."""
import hashlib
import os

import numpy as np
import pytest

from oracle import oracle_py as O


def load(golden_dir, name):
    return np.load(os.path.join(golden_dir, name))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("name", ["F1_step_random.npz", "F2_step_sigma_sources.npz"])
def test_step_bit_exact(golden_dir, name):
    z = load(golden_dir, name)
    I, J, K = (int(v) for v in z["dims"])
    g = O.Grid(I, J, K, float(z["disc"]))
    g.ep[:] = z["ep"]
    g.mu[:] = z["mu"]
    g.f[:] = z["fields0"]
    if "sig" in z:
        g.sig[:] = z["sig"]
        g.src[:] = z["sources"]
    for _ in range(int(z["steps"])):
        g.step(float(z["dt"]))
    assert np.array_equal(g.f, z["fields"]), np.abs(g.f - z["fields"]).max()


def dipole_run(z):
    N = int(z["dims"][0])
    g = O.Grid(N, N, N, float(z["disc"]))
    dt = float(z["dt"])
    c0 = 32 + N * (32 + N * 32)
    plane = N * N
    bufs = [np.zeros(12 * plane) for _ in z["omegas"]]
    view = O.Grid.__new__(O.Grid)
    for n in range(int(z["steps"])):
        g.src[6 * c0 + 2] = z["J"][n]
        g.step(dt)
        view_f = g.f[6 * plane * 48: 6 * plane * 49]
        view_s = O.OrcGrid(N, N, 1, g.disc, O.fptr(g.ep), O.fptr(g.mu), O.fptr(g.sig), O.dptr(view_f),
                           O.dptr(g.old), O.fptr(g.src))
        for q, w in enumerate(z["omegas"]):
            O.lib().orc_dft_volume(O.C.byref(view_s), n * dt * float(w), O.dptr(bufs[q]))
    return g, bufs


def test_dipole_mur_64_bit_exact(golden_dir):
    """F3 (64^3 dipole, Mur, 500 steps): full-field sha256 and the 5-wavelength Ez DFT on z=48."""
    z = load(golden_dir, "F3_dipole_mur_64.npz")
    g, bufs = dipole_run(z)
    N = g.I
    F = g.f.reshape(N, N, N, 6)
    assert np.array_equal(F[32, :, :, 2], z["ez_z32"])
    assert np.array_equal(F[48, :, :, 2], z["ez_z48"])
    assert sha(g.f) == str(z["fields_sha"])
    dft = np.stack([b.reshape(N, N, 6, 2)[:, :, 2, :] for b in bufs])
    assert np.array_equal(dft, z["dft_ez_z48"])


def test_mode_source_overlap_adjoint_bit_exact(golden_dir):
    """F4: soft mode source, y-slice DFT, three overlap monitors, then the adjoint run."""
    z = load(golden_dir, "F4_mode_source_overlap.npz")
    I, J, K = (int(v) for v in z["dims"])
    g = O.Grid(I, J, K, float(z["disc"]))
    g.ep[:] = z["ep"]
    g.mu[:] = z["mu"]
    dt, om = float(z["dt"]), float(z["omega"])
    m1 = z["mode1"].copy()
    m2 = z["mode2"].copy()
    mons = [(15, 0, m1, 30), (15, 1, m1, 30), (45, 0, m2, 24)]
    cums = [np.zeros(2) for _ in mons]
    F = np.zeros(12 * I * J * K)
    L = O.lib()
    for n in range(int(z["fwd_steps"])):
        amp = np.exp(-(n - 60) ** 2 / (2 * 25.0 ** 2))
        ph = n * dt * om
        L.orc_mode_source(g.ptr, I // 2, J // 2, 10, 30, 30, g.disc, O.dptr(m1), ph, amp, 0)
        g.step(dt)
        L.orc_dft_y_slice(g.ptr, ph, J // 2, O.dptr(F))
        for q, (kz, rev, mm, mi) in enumerate(mons):
            inst = np.zeros(2)
            L.orc_mode_overlap(g.ptr, I // 2, J // 2, kz, mi, mi, g.disc, O.dptr(mm), ph, rev, O.dptr(cums[q]),
                               O.dptr(inst))
            assert inst[0] == z["fwd_inst"][n, q, 0] and inst[1] == z["fwd_inst"][n, q, 1], (n, q)
    assert np.array_equal(np.array(cums), z["fwd_cum"])
    assert sha(g.f) == str(z["fwd_fields_sha"])
    assert np.array_equal(g.src.reshape(K, J, I, 6)[10], z["fwd_src_plane"])
    assert np.array_equal(F.reshape(K, J, I, 6, 2)[:, J // 2, :, 0:2, :], z["fwd_dft_y_exey"])
    assert np.array_equal(m1, z["mode1_after_fwd"])
    # adjoint: the rotated mode comes from the fixture (mode_smul is host complex arithmetic)
    m1 = z["mode1_rot"].copy()
    g.f[:] = 0
    g.old[:] = 0
    g.src[:] = 0
    for n in range(int(z["adj_steps"]), 0, -1):
        L.orc_mode_source_adjoint(g.ptr, I // 2, J // 2, 45, 30, 30, g.disc, O.dptr(m1), (n + 1) * dt * om, 0)
        g.step(dt)
    assert sha(g.f) == str(z["adj_fields_sha"])
    assert np.array_equal(g.src.reshape(K, J, I, 6)[45], z["adj_src_plane"])


def test_power_normalize_bit_exact(golden_dir):
    z = load(golden_dir, "F5_power_normalize.npz")
    mi, mj, md = int(z["mi"]), int(z["mj"]), float(z["mdisc"])
    m = z["mode0"].copy()
    p = np.zeros(2)
    O.lib().orc_mode_power_inner(mi, mj, md, O.dptr(m), O.dptr(m), O.dptr(p))
    assert np.array_equal(p, z["P0"])
    O.lib().orc_mode_power_normalize(mi, mj, md, O.dptr(m))
    assert np.array_equal(m, z["mode1"])
    O.lib().orc_mode_power_normalize(mi, mj, md, O.dptr(m))
    O.lib().orc_mode_power_normalize(mi, mj, md, O.dptr(m))
    assert np.array_equal(m, z["mode3"])
    # invariant: |P| = 1 after normalisation (to rounding)
    O.lib().orc_mode_power_inner(mi, mj, md, O.dptr(m), O.dptr(m), O.dptr(p))
    assert abs(np.hypot(*p) - 1.0) < 1e-13


def pattern_fields(T, n, a, b):
    idx = np.arange(n, dtype=np.int64)
    return np.stack([(((a * idx + b * t) % 97) / 48.5 - 1.0) for t in range(T)]).astype(np.float32)


def test_taper_bit_exact(golden_dir):
    """F6: edge store, ep-gradient, control-point projection, design updates, rasterisation."""
    z = load(golden_dir, "F6_taper.npz")
    I, J, K = (int(v) for v in z["dims"])
    seg, start, end, bottom, top, T = (int(z[k]) for k in ("seg", "start", "end", "bottom", "top", "T"))
    left, right = z["left"].astype(np.int32), z["right"].astype(np.int32)
    L = O.lib()
    g = O.Grid(I, J, K, 0.02)
    H, Ln = top - bottom, end - start
    hists = []
    for a, b in ((7, 13), (11, 5)):
        hist = np.zeros(6 * 2 * H * Ln * T)
        seq = pattern_fields(T, g.f.size, a, b)
        for t in range(T):
            g.f[:] = seq[t]
            L.orc_taper_store(g.ptr, seg, start, end, O.iptr(left), O.iptr(right), bottom, top, t, O.dptr(hist))
        hists.append(hist)
    assert np.array_equal(hists[0], z["fwd_hist"])
    assert np.array_equal(hists[1], z["adj_hist"])
    epg = np.zeros(2 * H * Ln)
    L.orc_taper_ep_gradient(H, Ln, T, O.dptr(hists[0]), O.dptr(hists[1]), O.dptr(epg))
    assert np.array_equal(epg, z["ep_grad"])
    lg, rg = np.zeros(seg + 1), np.zeros(seg + 1)
    L.orc_taper_grad_from_ep(seg, H, Ln, O.dptr(epg), float(z["ep_diff"]), O.dptr(lg), O.dptr(rg))
    assert np.array_equal(lg, z["left_grad"]) and np.array_equal(rg, z["right_grad"])
    for q, goal in enumerate(z["goals"]):
        l2, r2, lg2, rg2 = left.copy(), right.copy(), lg.copy(), rg.copy()
        fn = L.orc_taper_update_asymmetric if q < 2 else L.orc_taper_update_symmetric
        flag = fn(seg, O.iptr(l2), O.iptr(r2), O.dptr(lg2), O.dptr(rg2), float(goal), 0.02, 3)
        assert flag == z["upd_flag"][q]
        assert np.array_equal(l2, z["upd_left"][q]) and np.array_equal(r2, z["upd_right"][q])
        assert np.array_equal(lg2, z["upd_lg"][q]) and np.array_equal(rg2, z["upd_rg"][q])
    # rasterisation
    RI, RJ, RK = (int(v) for v in z["raster_dims"])
    g2 = O.Grid(RI, RJ, RK, 0.02)
    g2.ep[:] = np.float32(1.447 ** 2)
    ep4 = g2.ep.reshape(RK, RJ, RI, 3)
    mu4 = g2.mu.reshape(RK, RJ, RI, 3)
    ep4[:, 20:, :, :] = np.float32(1.93 ** 2)
    mu4[:, 20:, :, :] = np.float32(1.0)
    L.orc_set_taper(g2.ptr, 3.51 ** 2, 1.0, 20, 20, 14, 26, 0, 10, 12, 16)
    L.orc_taper_insert(g2.ptr, 3.51 ** 2, 1.0, seg, start, end, O.iptr(left), O.iptr(right), bottom, top)
    ep4[35:50, 12:16, 10:30, :] = np.float32(3.51 ** 2)
    assert np.array_equal(g2.ep, z["raster_ep"]) and np.array_equal(g2.mu, z["raster_mu"])


def test_complex_bit_exact(golden_dir):
    z = load(golden_dir, "F9_complex.npz")
    for x, want in zip(z["xs"], z["exp_imag"]):
        got = O.complex_exp(0.0, float(x))
        assert got.real == want[0] and got.imag == want[1]
    for (a, b), e, ang, sq in zip(z["zs"], z["exp_z"], z["angle"], z["sqrt"]):
        got = O.complex_exp(float(a), float(b))
        assert got.real == e[0] and got.imag == e[1]
        assert O.lib().orc_complex_angle(float(a), float(b)) == ang
        out = np.zeros(2)
        O.lib().orc_complex_sqrt(float(a), float(b), O.dptr(out))
        assert out[0] == sq[0] and out[1] == sq[1]
    # closed form: the Taylor exponential agrees with libm to a few ulp on [0, 2pi)
    for x in np.linspace(0, 2 * np.pi, 50):
        got = O.complex_exp(0.0, float(x))
        assert abs(got.real - np.cos(x)) < 1e-14 and abs(got.imag - np.sin(x)) < 1e-14
