"""Write the golden fixtures in tests/golden/ from the compiled reference.

Run in the build container (needs /root/reference to build oracle/_ref):
    make -C oracle ref && python tests/golden/make_golden.py          # every fixture
    python tests/golden/make_golden.py f11                            # only the named ones

Each fixture holds inputs and the reference's outputs for one hot-path function family; the
reference is called through oracle/ref_py.py.  Large outputs are stored as a sha256 of the
exact bytes plus slices for diagnostics.  manifest.json records parameters, seeds and the
sha256 of the reference sources the library was built from.
"""
import ctypes as C
import hashlib
import json
import os
import sys

import numpy as np

# The reference's z-slab Mur faces read plane k=1 / k=K-2 from the slab owner; with slabs thinner
# than 2 planes (K < 2*threads) that read races with a neighbour thread (timefield3d.c:2799-2848),
# so fixtures are generated with few threads and K >= 16.
os.environ["OMP_NUM_THREADS"] = "4"

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import ref_py as R  # noqa: E402

REF_DIR = "/root/reference"
DISC = 0.02
DT = 0.99 * DISC / np.sqrt(3)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def save(name, **arrays):
    np.savez_compressed(os.path.join(HERE, name), **arrays)


def new_tf(L, I, J, K):
    tf = L.timefield3d_new(I, J, K, DISC)
    return tf, R.tf_arrays(tf)


def f1_f2(L, manifest):
    """F1: random eps/mu/fields, 50 reference steps.  F2: adds random sigma and sources."""
    for name, (I, J, K), steps, lossy, seed in (("F1_step_random", (20, 14, 16), 50, False, 101),
                                                ("F2_step_sigma_sources", (14, 12, 16), 30, True, 202)):
        rng = np.random.default_rng(seed)
        tf, A = new_tf(L, I, J, K)
        A["ep"][:] = rng.uniform(1.0, 12.1, A["ep"].size).astype(np.float32)
        A["mu"][:] = rng.uniform(1.0, 2.0, A["mu"].size).astype(np.float32)
        f0 = rng.uniform(-1.0, 1.0, A["fields"].size).astype(np.float32)
        A["fields"][:] = f0
        extra = {}
        if lossy:
            A["sig"][:] = rng.uniform(0.0, 5.0, A["sig"].size).astype(np.float32)
            A["sources"][:] = rng.uniform(-1.0, 1.0, A["sources"].size).astype(np.float32)
            extra = {"sig": A["sig"].copy(), "sources": A["sources"].copy()}
        for _ in range(steps):
            L.timefield3d_apply_timestep_omp_full(tf, DT)
        save(name, dims=np.array([I, J, K]), disc=DISC, dt=DT, steps=steps, ep=A["ep"].copy(), mu=A["mu"].copy(),
             fields0=f0, fields=A["fields"].copy(), **extra)
        manifest[name] = {"dims": [I, J, K], "steps": steps, "seed": seed, "sigma_and_sources": lossy,
                          "calls": "timefield3d_apply_timestep_omp_full"}
        L.timefield3d_free(tf)


def dipole_J(n):
    """Gaussian-modulated sine, lambda0 = 1.55 um, bandwidth 0.2 um, t0 = 250 dt."""
    lam0, dlam = 1.55, 0.2
    w0 = 2 * np.pi / lam0
    sig_t = lam0 * lam0 / (2 * np.pi * dlam)
    t = n * DT
    t0 = 250 * DT
    return np.exp(-(t - t0) ** 2 / (2 * sig_t ** 2)) * np.sin(w0 * t)


def f3(L, manifest):
    """F3: 64^3 vacuum, Mur, z-polarised point dipole at the centre, 500 steps; Ez at the end and
    the DFT of Ez on plane z=48 at 5 wavelengths 1.50..1.60 um (reference DFT routine applied to a
    one-plane view of the grid, one fourier buffer per wavelength)."""
    N, steps = 64, 500
    tf, A = new_tf(L, N, N, N)
    c0 = 32 + N * (32 + N * 32)
    lams = np.linspace(1.50, 1.60, 5)
    omegas = 2 * np.pi / lams
    plane = N * N
    bufs = [np.zeros(12 * plane) for _ in lams]
    fields_ptr = C.cast(tf.fields, C.c_void_p).value
    Js = np.zeros(steps)
    for n in range(steps):
        Js[n] = np.float32(dipole_J(n))
        A["sources"][6 * c0 + 2] = Js[n]
        L.timefield3d_apply_timestep_omp_full(tf, DT)
        for q, w in enumerate(omegas):
            view = R.TimeField(N, N, 1, DISC, tf.ep, tf.mu, tf.sig,
                               C.cast(fields_ptr + 8 * 6 * plane * 48, C.POINTER(C.c_double)), tf.old_fields,
                               C.cast(bufs[q].ctypes.data, C.POINTER(R.Complex)), tf.sources)
            L.timefield3d_update_fourier_fields(view, n * DT * w)
    F = A["fields"].reshape(N, N, N, 6)
    dft_ez = np.stack([b.reshape(N, N, 6, 2)[:, :, 2, :] for b in bufs])  # [lam, y, x, re/im]
    save("F3_dipole_mur_64.npz", dims=np.array([N, N, N]), disc=DISC, dt=DT, steps=steps, J=Js,
         omegas=omegas, fields_sha=sha(A["fields"]), ez_z32=F[32, :, :, 2].copy(), ez_z48=F[48, :, :, 2].copy(),
         ez_sum_sq=np.sum(F[..., 2] ** 2), dft_ez_z48=dft_ez)
    manifest["F3_dipole_mur_64"] = {"dims": [N, N, N], "steps": steps, "source": "J_z at (32,32,32), float32",
                                   "lambdas_um": lams.tolist(), "t0_steps": 250,
                                   "phase": "omega * n * dt after step n"}
    L.timefield3d_free(tf)


def f4(L, manifest):
    """F4: Si strip, Gaussian modes from mode_gaussian; forward run with mode_source, y-slice DFT and
    three overlap monitors; then mode_smul + zero_fields + a time-reversed adjoint run."""
    I, J, K = 40, 40, 60
    om = 2 * np.pi / 1.55
    tf, A = new_tf(L, I, J, K)
    L.timefield3d_set_background(tf, 1.444 ** 2, 1.0)
    L.timefield3d_set_waveguide(tf, 3.48 ** 2, 14, 26, 0, K, 17, 23)
    md = L.mode_gaussian(30, 30, DISC, R.Complex(1.444 ** 2, 0), R.Complex(1, 0), om, 4.0, 0)
    md2 = L.mode_gaussian(24, 24, DISC, R.Complex(1.444 ** 2, 0), R.Complex(1, 0), om, 3.0, 1)
    mode1_0 = R.mode_fields(md).copy()
    mode2_0 = R.mode_fields(md2).copy()
    mons = [(15, 0, md), (15, 1, md), (45, 0, md2)]
    cums = [R.Complex(0, 0) for _ in mons]
    fwd_steps = 120
    inst = np.zeros((fwd_steps, len(mons), 2))
    for n in range(fwd_steps):
        amp = np.exp(-(n - 60) ** 2 / (2 * 25.0 ** 2))
        ph = n * DT * om
        L.timefield3d_mode_source(tf, I // 2, J // 2, 10, md, ph, amp, 0)
        L.timefield3d_apply_timestep_omp_full(tf, DT)
        L.timefield3d_update_fourier_fields_y_slice_omp(tf, ph, J // 2)
        for q, (z, rev, mm) in enumerate(mons):
            r = L.timefield3d_mode_overlap(tf, I // 2, J // 2, z, mm, ph, rev, C.byref(cums[q]))
            inst[n, q] = (r.re, r.im)
    FF = A["fourier"].reshape(K, J, I, 6, 2)[:, J // 2]
    fwd = dict(fwd_fields_sha=sha(A["fields"]), fwd_ex_plane=A["fields"].reshape(K, J, I, 6)[:, J // 2, :, 0].copy(),
               fwd_src_plane=A["sources"].reshape(K, J, I, 6)[10].copy(), fwd_inst=inst,
               fwd_cum=np.array([[c.re, c.im] for c in cums]), fwd_dft_y_exey=FF[:, :, 0:2, :].copy(),
               mode1_after_fwd=R.mode_fields(md).copy())
    L.mode_smul(R.Complex(0.3, 0.8), md)
    mode1_rot = R.mode_fields(md).copy()
    L.timefield3d_zero_fields(tf)
    adj_steps = 50
    for n in range(adj_steps, 0, -1):
        ph = (n + 1) * DT * om
        L.timefield3d_mode_source_adjoint(tf, I // 2, J // 2, 45, md, ph, 0)
        L.timefield3d_apply_timestep_omp_full(tf, DT)
    save("F4_mode_source_overlap.npz", dims=np.array([I, J, K]), disc=DISC, dt=DT, omega=om, ep=A["ep"].copy(),
         mu=A["mu"].copy(), mode1=mode1_0, mode2=mode2_0, mode1_rot=mode1_rot, fwd_steps=fwd_steps,
         adj_steps=adj_steps, adj_fields_sha=sha(A["fields"]),
         adj_hy_plane=A["fields"].reshape(K, J, I, 6)[:, J // 2, :, 4].copy(),
         adj_src_plane=A["sources"].reshape(K, J, I, 6)[45].copy(), **fwd)
    manifest["F4_mode_source_overlap"] = {
        "dims": [I, J, K], "background_eps": 1.444 ** 2, "core": "eps 3.48^2, x 14..26, y 17..23, all z",
        "mode1": "mode_gaussian(30,30,0.02,eps_bg,1,omega,4.0,0)", "mode2": "mode_gaussian(24,24,...,3.0,1)",
        "source": "plane z=10 centre (20,20), amplitude exp(-(n-60)^2/(2*25^2)), phase n*dt*omega",
        "monitors": "(z=15,rev 0,mode1), (z=15,rev 1,mode1), (z=45,rev 0,mode2)", "dft": "y slice j=20",
        "adjoint": "mode_smul(0.3+0.8i, mode1); zero_fields; n=50..1 source_adjoint z=45 phase (n+1)*dt*omega"}
    L.mode_free(md)
    L.mode_free(md2)
    L.timefield3d_free(tf)


def f5(L, manifest):
    """F5: mode_power_normalize applied 1x and 3x to a seeded random 7x5 mode."""
    rng = np.random.default_rng(505)
    md = L.mode_new(7, 5, 0.05)
    mf = R.mode_fields(md)
    mf[:] = rng.uniform(-1, 1, mf.size)
    m0 = mf.copy()
    p = L.mode_power_inner_product(md, md)
    L.mode_power_normalize(md)
    m1 = mf.copy()
    L.mode_power_normalize(md)
    L.mode_power_normalize(md)
    m3 = mf.copy()
    save("F5_power_normalize.npz", mi=7, mj=5, mdisc=0.05, mode0=m0, P0=np.array([p.re, p.im]), mode1=m1, mode3=m3)
    manifest["F5_power_normalize"] = {"mode": "7x5, disc 0.05, fields U[-1,1] seed 505"}
    L.mode_free(md)


def pattern_fields(T, n, a, b):
    """Deterministic field snapshots from integer arithmetic only: ((a*idx + b*t) % 97)/48.5 - 1."""
    idx = np.arange(n, dtype=np.int64)
    return np.stack([(((a * idx + b * t) % 97) / 48.5 - 1.0) for t in range(T)]).astype(np.float32)


def f6(L, manifest):
    """F6: taper edge store, ep-gradient, control-point projection, symmetric/asymmetric updates."""
    rng = np.random.default_rng(606)
    I, J, K = 30, 20, 40
    tf, A = new_tf(L, I, J, K)
    seg, start, end, bottom, top, T = 5, 5, 35, 8, 12, 20
    left = np.array([12, 11, 10, 9, 9, 8], np.int32)
    right = np.array([17, 18, 19, 21, 21, 22], np.int32)
    tp = L.adjoint3d_taper_new(3.51 ** 2, 1.0, seg, start, end, left.ctypes.data_as(C.POINTER(C.c_int)),
                               right.ctypes.data_as(C.POINTER(C.c_int)), bottom, top, T)
    H, Ln = top - bottom, end - start
    fields_seq = pattern_fields(T, A["fields"].size, 7, 13)
    for t in range(T):
        A["fields"][:] = fields_seq[t]
        L.adjoint3d_taper_store_forward_fields(tp, tf, t)
    adj_seq = pattern_fields(T, A["fields"].size, 11, 5)
    for t in range(T):
        A["fields"][:] = adj_seq[t]
        L.adjoint3d_taper_store_adjoint_fields(tp, tf, t)
    nh = 6 * 2 * H * Ln * T
    fwd = np.ctypeslib.as_array(tp.forward_fields, shape=(nh,)).copy()
    adj = np.ctypeslib.as_array(tp.adjoint_fields, shape=(nh,)).copy()
    L.adjoint3d_taper_calculate_ep_gradient(tp)
    epg = np.ctypeslib.as_array(tp.ep_gradient, shape=(2 * H * Ln,)).copy()
    ep_diff = 3.51 * 3.51 - 1.447 * 1.447
    L.adjoint3d_taper_calculate_gradients_from_ep_gradient(tp, ep_diff)
    lg = np.ctypeslib.as_array(tp.left_grad, shape=(seg + 1,)).copy()
    rg = np.ctypeslib.as_array(tp.right_grad, shape=(seg + 1,)).copy()
    updates = []
    for mode_name, goal in (("asymmetric", 0.05), ("asymmetric", 1e-7), ("symmetric", 0.05), ("symmetric", 1e-7)):
        t2 = L.adjoint3d_taper_new(3.51 ** 2, 1.0, seg, start, end, left.ctypes.data_as(C.POINTER(C.c_int)),
                                   right.ctypes.data_as(C.POINTER(C.c_int)), bottom, top, 1)
        np.ctypeslib.as_array(t2.left_grad, shape=(seg + 1,))[:] = lg
        np.ctypeslib.as_array(t2.right_grad, shape=(seg + 1,))[:] = rg
        fn = L.adjoint3d_taper_update_asymmetric if mode_name == "asymmetric" else L.adjoint3d_taper_update_symmetric
        flag = fn(t2, goal, DISC, 3)
        updates.append((flag, np.ctypeslib.as_array(t2.left, shape=(seg + 1,)).copy(),
                        np.ctypeslib.as_array(t2.right, shape=(seg + 1,)).copy(),
                        np.ctypeslib.as_array(t2.left_grad, shape=(seg + 1,)).copy(),
                        np.ctypeslib.as_array(t2.right_grad, shape=(seg + 1,)).copy()))
    # rasterisation of the taper into a background + a waveguide
    tf2, A2 = new_tf(L, 40, 30, 50)
    L.timefield3d_set_background(tf2, 1.447 ** 2, 1.0)
    L.timefield3d_set_top_half_plane(tf2, 1.93 ** 2, 1.0, 20)
    L.timefield3d_set_taper(tf2, 3.51 ** 2, 1.0, 20, 20, 14, 26, 0, 10, 12, 16)
    L.adjoint3d_taper_insert(tf2, tp)
    L.timefield3d_set_waveguide(tf2, 3.51 ** 2, 10, 30, 35, 50, 12, 16)
    save("F6_taper.npz", dims=np.array([I, J, K]), seg=seg, start=start, end=end, bottom=bottom, top=top, T=T,
         left=left, right=right, fwd_hist=fwd, adj_hist=adj, ep_grad=epg,
         ep_diff=ep_diff, left_grad=lg, right_grad=rg, goals=np.array([0.05, 1e-7, 0.05, 1e-7]),
         upd_flag=np.array([u[0] for u in updates]), upd_left=np.stack([u[1] for u in updates]),
         upd_right=np.stack([u[2] for u in updates]), upd_lg=np.stack([u[3] for u in updates]),
         upd_rg=np.stack([u[4] for u in updates]), raster_dims=np.array([40, 30, 50]), raster_ep=A2["ep"].copy(),
         raster_mu=A2["mu"].copy())
    manifest["F6_taper"] = {"grid": [I, J, K], "seg": seg, "start": start, "end": end, "bottom": bottom,
                            "top": top, "T": T, "min_width": 3,
                            "fields": "pattern_fields(T, 6n, 7, 13) forward, (11, 5) adjoint",
                            "updates": "asym 0.05, asym 1e-7, sym 0.05, sym 1e-7 on copies",
                            "raster": "background 1.447^2; top half plane 1.93^2 from y=20; "
                                      "set_taper(3.51^2,1,20,20,14,26,0,10,12,16); insert; "
                                      "waveguide(3.51^2,10,30,35,50,12,16)"}
    L.timefield3d_free(tf)
    L.timefield3d_free(tf2)


def f7(L, manifest):
    """F7: the explicit-conductivity serial step (timefield3d_apply_timestep), and the 3-D block
    adjoint (store, reverse-time gradient accumulation, permittivity adjustment, taper projection)."""
    rng = np.random.default_rng(707)
    I, J, K = 14, 12, 16
    tf, A = new_tf(L, I, J, K)
    A["ep"][:] = rng.uniform(1.0, 12.1, A["ep"].size).astype(np.float32)
    A["mu"][:] = rng.uniform(1.0, 2.0, A["mu"].size).astype(np.float32)
    A["sig"][:] = rng.uniform(0.0, 3.0, A["sig"].size).astype(np.float32)
    A["sources"][:] = rng.uniform(-1.0, 1.0, A["sources"].size).astype(np.float32)
    f0 = rng.uniform(-1.0, 1.0, A["fields"].size).astype(np.float32)
    A["fields"][:] = f0
    ex_in = {k: A[k].copy() for k in ("ep", "mu", "sig", "sources")}
    for _ in range(20):
        L.timefield3d_apply_timestep(tf, DT)
    ex_out = A["fields"].copy()
    L.timefield3d_free(tf)
    # block adjoint
    BI, BJ, BK, T = 20, 16, 18, 6
    box = (4, 12, 3, 9, 5, 14)
    tf, A = new_tf(L, BI, BJ, BK)
    blk = L.ajoint3d_new(*box, T)
    fwd = pattern_fields(T, A["fields"].size, 7, 13)
    adj = pattern_fields(T, A["fields"].size, 11, 5)
    for t in range(T):
        A["fields"][:] = fwd[t]
        L.adjoint3d_store_fwd_fields(blk, tf, t)
    for t in range(T - 1, -1, -1):
        A["fields"][:] = adj[t]
        L.adjoint3d_update_gradient(blk, tf, t)
    nb = (box[1] - box[0]) * (box[3] - box[2]) * (box[5] - box[4])
    grad = np.ctypeslib.as_array(blk.gradient, shape=(nb,)).copy()
    L.timefield3d_set_background(tf, 2.0, 1.0)
    L.adjoint3d_adjust_permittivity(blk, tf, 4.0, 0.05, 1.5, 12.0)
    ep_after = A["ep"].copy()
    # projection of the block gradient onto a taper inside the box
    left = np.array([6, 6, 5, 5], np.int32)
    right = np.array([9, 9, 10, 10], np.int32)
    tp = L.adjoint3d_taper_new(12.0, 1.0, 3, 5, 14, left.ctypes.data_as(C.POINTER(C.c_int)),
                               right.ctypes.data_as(C.POINTER(C.c_int)), 3, 9, 1)
    L.adjoint3d_taper_calculate_gradients(tp, blk, 10.0)
    tlg = np.ctypeslib.as_array(tp.left_grad, shape=(4,)).copy()
    trg = np.ctypeslib.as_array(tp.right_grad, shape=(4,)).copy()
    save("F7_explicit_block.npz", dims=np.array([I, J, K]), dt=DT, disc=DISC, steps=20, fields0=f0, fields=ex_out,
         block_dims=np.array([BI, BJ, BK]), box=np.array(box), T=T, grad=grad, ep_after=ep_after,
         taper_left=left, taper_right=right, taper_left_grad=tlg, taper_right_grad=trg, **ex_in)
    manifest["F7_explicit_block"] = {
        "explicit": "14x12x16 random eps/mu/sigma/sources/fields, 20 x timefield3d_apply_timestep",
        "block": "20x16x18 grid, box (4,12,3,9,5,14), T=6, fields pattern_fields(7,13) forward, (11,5) reverse; "
                 "adjust_permittivity(omega 4, improvement 0.05, [1.5, 12]) on background 2; taper seg 3 "
                 "z 5..14 y 3..9 left [6,6,5,5] right [9,9,10,10] calculate_gradients(ep_diff 10)"}
    L.timefield3d_free(tf)


def f9(L, manifest):
    """F9: complex_exp (Taylor series) on the imaginary axis and off it, complex_angle, complex_sqrt."""
    rng = np.random.default_rng(909)
    xs = np.concatenate([rng.uniform(0, 2 * np.pi, 150), rng.uniform(-20, 20, 50)])
    zs = rng.uniform(-3, 3, (60, 2))
    e_im = np.array([[r.re, r.im] for r in (L.complex_exp(R.Complex(0, x)) for x in xs)])
    e_z = np.array([[r.re, r.im] for r in (L.complex_exp(R.Complex(a, b)) for a, b in zs)])
    ang = np.array([L.complex_angle(R.Complex(a, b)) for a, b in zs])
    sq = np.array([[r.re, r.im] for r in (L.complex_sqrt(R.Complex(a, b)) for a, b in zs)])
    save("F9_complex.npz", xs=xs, exp_imag=e_im, zs=zs, exp_z=e_z, angle=ang, sqrt=sq)
    manifest["F9_complex"] = {"inputs": "150 U[0,2pi) + 50 U[-20,20) imaginary; 60 U[-3,3)^2 complex"}


def f11(L, manifest):
    """F11: the mode solver's transverse-H operator (mode.c:2796-2943) on a Si strip cross-section
    (n 3.48 in 1.44, 20 nm cells, 30 x 24) with an order-2 PML of 4 cells (mode.c:1634) and
    random complex Hx, Hy (all six components set, as the operator must ignore the others)."""
    rng = np.random.default_rng(1111)
    mi, mj, disc, omega = 30, 24, DISC, 2 * np.pi / 1.55
    md = L.mode_new(mi, mj, disc)
    n = mi * mj
    ep = np.full((mj, mi, 3), 1.44 ** 2, complex)
    ep[9:15, 6:24, :] = 3.48 ** 2
    mu = np.ones((mj, mi, 3), complex)
    mu[:, :, 2] += 0.05j * rng.uniform(size=(mj, mi))  # exercises the mu_z division
    fields = rng.normal(size=(n * 6,)) + 1j * rng.normal(size=(n * 6,))

    def put(mat, arr):
        v = R.view(mat.m, arr.size, np.float64).reshape(-1, 2)
        v[:, 0] = arr.reshape(-1).real
        v[:, 1] = arr.reshape(-1).imag

    put(md.ep, ep)
    put(md.mu, mu)
    put(md.fields, fields)
    L.mode_set_pml(md, 0.8, 2, 4)
    pml = R.view(md.pml.m, 2 * n, np.float64).copy()
    out = L.mode_apply_beta_squared_operator(md, omega)
    res = R.view(out.fields.m, 6 * n, np.float64).copy()
    L.mode_free(out)
    L.mode_free(md)
    save("F11_mode_operator.npz", dims=np.array([mi, mj]), disc=disc, omega=omega, ep=ep, mu=mu, fields=fields,
         pml=pml, out=res)
    manifest["F11_mode_operator"] = {"grid": [mi, mj], "disc_um": disc, "lambda_um": 1.55,
                                     "core": "rows 9-14, columns 6-23, eps 3.48^2 in 1.44^2",
                                     "pml": "mode_set_pml(0.8, order 2, 4 cells)", "seed": 1111}


def f10(L, manifest):
    """F10: the mode text file (mode.c:843-1038).  A 9 x 7 mode with random complex ep, mu and
    fields (magnitudes up to 1e3, both signs) written by the reference with dimension 2 and 1, and
    the reference's own loads of those two files."""
    import tempfile
    rng = np.random.default_rng(1010)
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    mi, mj, mdisc = 9, 7, 0.0125
    md = L.mode_new(mi, mj, mdisc)
    arrays = {}
    for name, mat, n in (("ep", md.ep, 3), ("mu", md.mu, 3), ("fields", md.fields, 6)):
        v = R.view(mat.m, n * mi * mj, np.float64)  # n*mi*mj complex entries as interleaved doubles
        v[:] = rng.uniform(-1, 1, v.size) * 10.0 ** rng.integers(-3, 4, v.size)
        arrays[name] = v.copy()
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for dim in (2, 1):
            path = os.path.join(tmp, f"mode{dim}.txt").encode()
            f = libc.fopen(path, b"w")
            L.mode_file_print(f, md, dim)
            libc.fclose(f)
            out[f"file{dim}"] = np.frombuffer(open(path, "rb").read(), np.uint8).copy()
            f = libc.fopen(path, b"r")
            ld = L.mode_file_load(f)
            libc.fclose(f)
            out[f"load{dim}_ep"] = R.view(ld.ep.m, 3 * mi * mj, np.float64).copy()
            out[f"load{dim}_mu"] = R.view(ld.mu.m, 3 * mi * mj, np.float64).copy()
            out[f"load{dim}_fields"] = R.view(ld.fields.m, 6 * mi * mj, np.float64).copy()
            L.mode_free(ld)
    L.mode_free(md)
    save("F10_mode_file.npz", mi=mi, mj=mj, mdisc=mdisc, **{"in_" + k: v for k, v in arrays.items()}, **out)
    manifest["F10_mode_file"] = {"mode": "9x7, disc 0.0125, ep/mu/fields U[-1,1] x 10^U{-3..3} seed 1010",
                                 "files": "dimension 2 and dimension 1, with the reference's loads of each"}


def f13(L, manifest):
    """F13: the Gaussian TE and TM sources (timefield3d.c:3070-3174) on a 30 x 26 x 12 grid with
    random ep, mu and sigma: the whole source array after a TE source of width 0.2 um at
    (15, 13, 5) (its 30 x 30 mode window wraps across rows) and a TM source of width 0.1 um at
    (4, 20, 8) near the corner."""
    rng = np.random.default_rng(1313)
    I, J, K = 30, 26, 12
    tf, A = new_tf(L, I, J, K)
    A["ep"][:] = rng.uniform(1.0, 12.0, A["ep"].size).astype(np.float32)
    A["mu"][:] = rng.uniform(1.0, 2.0, A["mu"].size).astype(np.float32)
    A["sig"][:] = rng.uniform(0.0, 0.5, A["sig"].size).astype(np.float32)
    ep, mu, sig = A["ep"].copy(), A["mu"].copy(), A["sig"].copy()
    omega = 2 * np.pi / 1.55
    L.timefield3d_TE_gaussian_source(tf, 15, 13, 5, 0.2, omega, 0.7, 1.3)
    L.timefield3d_TM_gaussian_source(tf, 4, 20, 8, 0.1, omega, -2.1, 0.6)
    save("F13_gaussian_sources.npz", dims=np.array([I, J, K]), disc=DISC, ep=ep, mu=mu, sig=sig, omega=omega,
         te=np.array([15, 13, 5, 0.2, 0.7, 1.3]), tm=np.array([4, 20, 8, 0.1, -2.1, 0.6]),
         sources=A["sources"].copy())
    manifest["F13_gaussian_sources"] = {"grid": [I, J, K], "materials": "ep U[1,12], mu U[1,2], sig U[0,0.5] seed 1313",
                                        "TE": "(15,13,5) width 0.2 phase 0.7 amplitude 1.3",
                                        "TM": "(4,20,8) width 0.1 phase -2.1 amplitude 0.6", "lambda_um": 1.55}
    L.timefield3d_free(tf)


def f12(L, manifest):
    """F12: the general matrix and mode helpers -- every minor of a 4 x 5 matrix, determinants of
    1 x 1 to 6 x 6 random and of a singular 5 x 5, the closed-form 2 x 2 eigenvalues, the QL
    eigenvalue sweep on matrices with eigenvalue magnitudes 1..n, and the plain mode inner
    product, norm and normalisation of a 7 x 5 mode and its zero-gradient edges (matrix.c:162-240,
    873-1111; mode.c:1208-1350, 1484-1530)."""
    rng = np.random.default_rng(1212)

    def rc(*shape):
        return rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)

    def call(fn, A, *args):
        M = R.matrix_from(L, A)
        out = fn(M, *args)
        L.matrix_free(M)
        return out

    sub_in = rc(4, 5)
    subs = []
    for i in range(4):
        for j in range(5):
            out = call(L.matrix_get_submatrix, sub_in, i, j)
            subs.append(R.matrix_array(out))
            L.matrix_free(out)
    det_in = [rc(n, n) for n in range(1, 7)]
    sing = rc(5, 5)
    sing[3] = sing[1] * (0.5 - 0.25j)  # row 3 is a multiple of row 1
    det_in.append(sing)
    dets = []
    for A in det_in:
        z = call(L.matrix_determinant, A)
        dets.append(z.re + 1j * z.im)
    two_in = rc(20, 2, 2)
    two_in[5] = [[1, 2], [0, 1]]         # defective: one double root
    two_in[6] = [[0, -1], [1, 0]]        # roots +-i
    two_out = []
    for A in two_in:
        out = call(L.matrix_diagonalize_2_by_2, A)
        two_out.append(np.diag(R.matrix_array(out)))
        L.matrix_free(out)
    nh_in, nh_out = [], []
    for n in (2, 3, 4, 6):
        lam = np.arange(1, n + 1) * np.exp(1j * rng.uniform(0, 2 * np.pi, n))
        V = rc(n, n) + 3 * np.eye(n)
        A = V @ np.diag(lam) @ np.linalg.inv(V)
        out = call(L.matrix_diagonalize_non_hermitian, A, 1e-12, 400)
        nh_in.append(A)
        nh_out.append(R.matrix_array(out))
        L.matrix_free(out)
    mi, mj, mdisc = 7, 5, 0.05
    mds = []
    for _ in range(2):
        md = L.mode_new(mi, mj, mdisc)
        mf = R.mode_fields(md)
        mf[:] = rng.uniform(-1, 1, mf.size)
        mds.append(md)
    psi0, phi0 = R.mode_fields(mds[0]).copy(), R.mode_fields(mds[1]).copy()
    ip = L.mode_inner_product(mds[0], mds[1])
    nrm = L.mode_norm(mds[0])
    L.mode_normalize(mds[0])
    psi1 = R.mode_fields(mds[0]).copy()
    nrm1 = L.mode_norm(mds[0])
    L.mode_apply_zero_gradient_bc(mds[1])
    phi_bc = R.mode_fields(mds[1]).copy()
    for md in mds:
        L.mode_free(md)
    save("F12_matrix_mode.npz", phi_bc=phi_bc, sub_in=sub_in, subs=np.array(subs), det_sizes=np.array([len(A) for A in det_in]),
         det_in=np.concatenate([A.reshape(-1) for A in det_in]), dets=np.array(dets), two_in=two_in,
         two_out=np.array(two_out), nh_sizes=np.array([len(A) for A in nh_in]),
         nh_in=np.concatenate([A.reshape(-1) for A in nh_in]),
         nh_out=np.concatenate([A.reshape(-1) for A in nh_out]), nh_tol=1e-12, nh_max_iterations=400,
         mi=mi, mj=mj, mdisc=mdisc, psi0=psi0, phi0=phi0, inner=np.array([ip.re, ip.im]), norm=nrm,
         psi1=psi1, norm1=nrm1)
    manifest["F12_matrix_mode"] = {
        "matrix": "U[-1,1]^2 complex entries seed 1212: minors of 4x5; det 1..6 and singular 5x5; 20 2x2 "
                  "(one defective, one rotation); QL sweep (tol 1e-12, 400 sweeps) on V diag(1..n e^{i phi}) V^-1, "
                  "n 2,3,4,6",
        "mode": "two 7x5 modes, disc 0.05, fields U[-1,1]: inner product, norm, normalize, norm after; "
                "zero-gradient edges of the second"}


def main():
    L = R.lib()
    only = sys.argv[1:]  # fixture function names (e.g. f11): regenerate just those
    if only:
        with open(os.path.join(HERE, "manifest.json")) as fh:
            manifest = json.load(fh)
        for name in only:
            globals()[name](L, manifest)
        with open(os.path.join(HERE, "manifest.json"), "w") as fh:
            json.dump(manifest, fh, indent=1, sort_keys=True)
        print("fixtures", only, "written to", HERE)
        return
    manifest = {"reference_sources_sha256": {}, "compiler": "gcc -O3 -fopenmp -fPIC (oracle/Makefile ref)",
                "disc_um": DISC, "dt": DT}
    for fn in ("complex.c", "matrix.c", "mode.c", "timefield3d.c", "adjoint3d.c"):
        with open(os.path.join(REF_DIR, fn), "rb") as fh:
            manifest["reference_sources_sha256"][fn] = hashlib.sha256(fh.read()).hexdigest()
    f1_f2(L, manifest)
    f3(L, manifest)
    f4(L, manifest)
    f5(L, manifest)
    f6(L, manifest)
    f7(L, manifest)
    f9(L, manifest)
    f10(L, manifest)
    f11(L, manifest)
    f12(L, manifest)
    f13(L, manifest)
    with open(os.path.join(HERE, "manifest.json"), "w") as fh:
        json.dump(manifest, fh, indent=1, sort_keys=True)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
