"""CPU-side checks of libfdtd3d.so: it loads, exports every function include/*.h declares, its
host-only functions match the reference (fixtures), and the reference driver links against it.
No GPU call is made (device state is created lazily on the first device operation)."""
import ctypes as C
import glob
import importlib
import os
import re
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

fd = importlib.import_module("3d_fdtd_amd")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(INC, "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", " ", src, flags=re.S)
        src = re.sub(r"//[^\n]*", " ", src)
        src = re.sub(r"#[^\n]*", " ", src)
        for m in re.finditer(r"\b([A-Za-z_]\w*)\s*\(([^;{}()]|\([^()]*\))*\)\s*;", src):
            name = m.group(1)
            if name in ("if", "while", "return", "sizeof"):
                continue
            names.add(name)
    return names


def test_library_exports_every_declared_function():
    lib = C.CDLL(fd.LIB_PATH)
    names = declared_functions()
    assert len(names) > 150
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_product_complex_matches_reference_fixture(golden_dir):
    z = np.load(os.path.join(golden_dir, "F9_complex.npz"))
    L = fd.lib()
    for x, want in zip(z["xs"], z["exp_imag"]):
        r = L.complex_exp(fd.Complex(0.0, float(x)))
        assert r.re == want[0] and r.im == want[1]
    for (a, b), e, ang, sq in zip(z["zs"], z["exp_z"], z["angle"], z["sqrt"]):
        r = L.complex_exp(fd.Complex(float(a), float(b)))
        assert r.re == e[0] and r.im == e[1]
        assert L.complex_angle(fd.Complex(float(a), float(b))) == ang
        s = L.complex_sqrt(fd.Complex(float(a), float(b)))
        assert s.re == sq[0] and s.im == sq[1]


def test_product_modes_match_reference_fixture(golden_dir):
    z = np.load(os.path.join(golden_dir, "F4_mode_source_overlap.npz"))
    L = fd.lib()
    om = float(z["omega"])
    m1 = L.mode_gaussian(30, 30, 0.02, fd.Complex(1.444 ** 2, 0), fd.Complex(1, 0), om, 4.0, 0)
    m2 = L.mode_gaussian(24, 24, 0.02, fd.Complex(1.444 ** 2, 0), fd.Complex(1, 0), om, 3.0, 1)
    assert np.array_equal(fd.mode_fields(m1), z["mode1"])
    assert np.array_equal(fd.mode_fields(m2), z["mode2"])
    L.mode_smul(fd.Complex(0.3, 0.8), m1)
    # the fixture rotated the mode after 120 source/overlap renormalisations; check the rotation alone
    ref = z["mode1"][0::2] + 1j * z["mode1"][1::2]
    got = fd.mode_fields(m1)[0::2] + 1j * fd.mode_fields(m1)[1::2]
    assert np.allclose(got, ref * (0.3 + 0.8j), rtol=0, atol=1e-15)
    z5 = np.load(os.path.join(golden_dir, "F5_power_normalize.npz"))
    m = fd.mode_from_array(int(z5["mi"]), int(z5["mj"]), float(z5["mdisc"]), z5["mode0"])
    p = L.mode_power_inner_product(m, m)
    assert p.re == z5["P0"][0] and p.im == z5["P0"][1]
    L.mode_power_normalize(m)
    assert np.array_equal(fd.mode_fields(m), z5["mode1"])
    L.mode_power_normalize(m)
    L.mode_power_normalize(m)
    assert np.array_equal(fd.mode_fields(m), z5["mode3"])


def test_product_taper_host_logic_matches_reference_fixture(golden_dir):
    z = np.load(os.path.join(golden_dir, "F6_taper.npz"))
    L = fd.lib()
    seg, start, end, bottom, top = (int(z[k]) for k in ("seg", "start", "end", "bottom", "top"))
    ip = C.POINTER(C.c_int)
    left = np.ascontiguousarray(z["left"].astype(np.int32))
    right = np.ascontiguousarray(z["right"].astype(np.int32))
    # projection of the ep-gradient onto control points, then the design updates
    for q, goal in enumerate(z["goals"]):
        tp = L.adjoint3d_taper_new(3.51 ** 2, 1.0, seg, start, end, left.ctypes.data_as(ip), right.ctypes.data_as(ip),
                                   bottom, top, 1)
        n = 2 * (top - bottom) * (end - start)
        np.ctypeslib.as_array(tp.ep_gradient, shape=(n,))[:] = z["ep_grad"]
        L.adjoint3d_taper_calculate_gradients_from_ep_gradient(tp, float(z["ep_diff"]))
        lg = np.ctypeslib.as_array(tp.left_grad, shape=(seg + 1,))
        rg = np.ctypeslib.as_array(tp.right_grad, shape=(seg + 1,))
        assert np.array_equal(lg, z["left_grad"]) and np.array_equal(rg, z["right_grad"])
        fn = L.adjoint3d_taper_update_asymmetric if q < 2 else L.adjoint3d_taper_update_symmetric
        flag = fn(tp, float(goal), 0.02, 3)
        assert flag == z["upd_flag"][q]
        assert np.array_equal(np.ctypeslib.as_array(tp.left, shape=(seg + 1,)), z["upd_left"][q])
        assert np.array_equal(np.ctypeslib.as_array(tp.right, shape=(seg + 1,)), z["upd_right"][q])
    # rasterisation through the product's setters (host side, no device state created)
    RI, RJ, RK = (int(v) for v in z["raster_dims"])
    g = fd.Grid(RI, RJ, RK, 0.02)
    L.timefield3d_set_background(g.tf, 1.447 ** 2, 1.0)
    L.timefield3d_set_top_half_plane(g.tf, 1.93 ** 2, 1.0, 20)
    L.timefield3d_set_taper(g.tf, 3.51 ** 2, 1.0, 20, 20, 14, 26, 0, 10, 12, 16)
    tp = L.adjoint3d_taper_new(3.51 ** 2, 1.0, seg, start, end, left.ctypes.data_as(ip), right.ctypes.data_as(ip),
                               bottom, top, 1)
    L.adjoint3d_taper_insert(g.tf, tp)
    L.timefield3d_set_waveguide(g.tf, 3.51 ** 2, 10, 30, 35, 50, 12, 16)
    assert np.array_equal(g.a["ep"], z["raster_ep"]) and np.array_equal(g.a["mu"], z["raster_mu"])
    g.free()


def test_mode_file_round_trip_and_bad_file(tmp_path):
    L = fd.lib()
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    L.mode_file_print.argtypes = [C.c_void_p, fd.Mode, C.c_int]
    L.mode_file_load.argtypes = [C.c_void_p]
    L.mode_file_load.restype = fd.Mode
    m = L.mode_gaussian(12, 10, 0.02, fd.Complex(2.0, 0), fd.Complex(1, 0), 4.0, 2.0, 0)
    path = str(tmp_path / "m.txt").encode()
    f = libc.fopen(path, b"w")
    L.mode_file_print(f, m, 2)
    libc.fclose(f)
    f = libc.fopen(path, b"r")
    r = L.mode_file_load(f)
    libc.fclose(f)
    assert (r.i, r.j) == (12, 10)
    # the text format keeps 6 decimals
    assert np.allclose(fd.mode_fields(r), fd.mode_fields(m), atol=6e-7)
    # a truncated file aborts with a message instead of reading garbage (checked in a child)
    trunc = tmp_path / "t.txt"
    trunc.write_text(open(path.decode()).read()[:200])
    code = ("import ctypes as C, importlib, sys; sys.path.insert(0, %r); fd = importlib.import_module('3d_fdtd_amd');"
            "L = fd.lib(); libc = C.CDLL(None); libc.fopen.restype = C.c_void_p;"
            "libc.fopen.argtypes = [C.c_char_p, C.c_char_p]; L.mode_file_load.argtypes = [C.c_void_p];"
            "L.mode_file_load.restype = fd.Mode; L.mode_file_load(libc.fopen(%r, b'r'))" % (ROOT, str(trunc).encode()))
    r = subprocess.run(["python", "-c", code], capture_output=True, text=True)
    assert r.returncode != 0 and "truncated" in r.stderr


@pytest.mark.skipif(not os.path.exists("/root/reference/3d_fdtd.c"), reason="reference driver not present")
def test_reference_driver_links_unchanged():
    d = tempfile.mkdtemp()
    try:
        shutil.copy("/root/reference/3d_fdtd.c", d)
        r = subprocess.run(["gcc", "-O2", "-fopenmp", "-I" + INC, os.path.join(d, "3d_fdtd.c"), "-o",
                            os.path.join(d, "drv"), "-L" + os.path.dirname(fd.LIB_PATH), "-lfdtd3d", "-lm"],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
    finally:
        shutil.rmtree(d)
