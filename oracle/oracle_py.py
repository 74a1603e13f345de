"""ctypes binding of the clean-room FP64 oracle (oracle/build/liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py use this module, and
only as the checker.  Arrays use the reference AoS layout (see fdtd_oracle.h).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "build", "liboracle.so")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


class OrcGrid(C.Structure):
    _fields_ = [("I", C.c_int), ("J", C.c_int), ("K", C.c_int), ("disc", C.c_double), ("ep", _fp), ("mu", _fp),
                ("sig", _fp), ("f", _dp), ("old", _dp), ("src", _fp)]


class OrcCpmlParams(C.Structure):
    _fields_ = [("npml", C.c_int * 6), ("order", C.c_int), ("R0", C.c_double), ("kappa_max", C.c_double),
                ("alpha_max", C.c_double)]


class OrcCpml(C.Structure):
    _fields_ = [("nlo", C.c_int * 3), ("nhi", C.c_int * 3),
                ("bE", _dp * 3), ("aE", _dp * 3), ("kE", _dp * 3),
                ("bH", _dp * 3), ("aH", _dp * 3), ("kH", _dp * 3),
                ("psi", _dp * 18)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_LIB):
        build()
    L = C.CDLL(ORACLE_LIB)
    G = C.POINTER(OrcGrid)
    d, i = C.c_double, C.c_int
    sig = {
        "orc_complex_exp": ([d, d, _dp], None),
        "orc_complex_angle": ([d, d], d),
        "orc_complex_sqrt": ([d, d, _dp], None),
        "orc_reduce_phase": ([d], d),
        "orc_snapshot_shells": ([G], None),
        "orc_update_E": ([G, d], None),
        "orc_update_H": ([G, d], None),
        "orc_mur_E": ([G, d], None),
        "orc_mur_H": ([G, d], None),
        "orc_step": ([G, d], None),
        "orc_dft_volume": ([G, d, _dp], None),
        "orc_dft_y_slice": ([G, d, i, _dp], None),
        "orc_mode_power_inner": ([i, i, d, _dp, _dp, _dp], None),
        "orc_mode_power_normalize": ([i, i, d, _dp], None),
        "orc_mode_source": ([G, i, i, i, i, i, d, _dp, d, d, i], None),
        "orc_mode_source_adjoint": ([G, i, i, i, i, i, d, _dp, d, i], None),
        "orc_mode_overlap": ([G, i, i, i, i, i, d, _dp, d, i, _dp, _dp], None),
        "orc_set_taper": ([G, d, d, i, i, i, i, i, i, i, i], None),
        "orc_taper_insert": ([G, d, d, i, i, i, _ip, _ip, i, i], None),
        "orc_taper_store": ([G, i, i, i, _ip, _ip, i, i, i, _dp], None),
        "orc_taper_ep_gradient": ([i, i, i, _dp, _dp, _dp], None),
        "orc_taper_grad_from_ep": ([i, i, i, _dp, d, _dp, _dp], None),
        "orc_taper_update_asymmetric": ([i, _ip, _ip, _dp, _dp, d, d, i], i),
        "orc_taper_update_symmetric": ([i, _ip, _ip, _dp, _dp, d, d, i], i),
        "orc_cpml_profile": ([i, i, i, C.POINTER(OrcCpmlParams), d, d, _dp, _dp, _dp, _dp, _dp, _dp], None),
        "orc_update_E_cpml": ([G, C.POINTER(OrcCpml), d], None),
        "orc_update_H_cpml": ([G, C.POINTER(OrcCpml), d], None),
        "orc_step_cpml": ([G, C.POINTER(OrcCpml), d], None),
        "orc_dft_scalar": ([i, _dp, d, d, _dp], None),
        "orc_dft_box": ([G, _ip, i, i, _dp, d, _dp], None),
        "orc_overlap_adjoint_source": ([G, i, i, i, i, i, d, _dp, i, d, d, d], None),
        "orc_dft_gradient": ([C.c_long, i, i, _dp, i, i, _dp, i, d, d, _dp], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def dptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def fptr(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_fp)


def iptr(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


class Grid:
    """Owns reference-layout arrays and the OrcGrid struct that points at them."""

    def __init__(self, I, J, K, disc):
        n = I * J * K
        self.I, self.J, self.K, self.disc = I, J, K, disc
        self.ep = np.ones(3 * n, np.float32)
        self.mu = np.ones(3 * n, np.float32)
        self.sig = np.zeros(6 * n, np.float32)
        self.f = np.zeros(6 * n, np.float64)
        self.old = np.zeros(6 * n, np.float64)
        self.src = np.zeros(6 * n, np.float32)
        self._refresh()

    def _refresh(self):
        self.s = OrcGrid(self.I, self.J, self.K, self.disc, fptr(self.ep), fptr(self.mu), fptr(self.sig),
                         dptr(self.f), dptr(self.old), fptr(self.src))

    @property
    def ptr(self):
        return C.byref(self.s)

    def comp(self, m):
        """Component m (0..5) as a [K, J, I] view."""
        return self.f.reshape(self.K, self.J, self.I, 6)[..., m]

    def step(self, dt):
        lib().orc_step(self.ptr, dt)


class Cpml:
    """Profile tables + full-volume psi arrays for orc_*_cpml."""

    def __init__(self, grid, npml, dt, order=3, R0=1e-8, kappa_max=1.0, alpha_max=0.0):
        N = (grid.I, grid.J, grid.K)
        n = grid.I * grid.J * grid.K
        p = OrcCpmlParams((C.c_int * 6)(*npml), order, R0, kappa_max, alpha_max)
        self.params = p
        self.tables = {}
        s = OrcCpml()
        for a in range(3):
            s.nlo[a] = npml[2 * a]
            s.nhi[a] = npml[2 * a + 1]
            tabs = [np.zeros(N[a], np.float64) for _ in range(6)]
            self.tables[a] = tabs
            lib().orc_cpml_profile(N[a], npml[2 * a], npml[2 * a + 1], C.byref(p), grid.disc, dt,
                                   *[dptr(t) for t in tabs])
            s.bE[a], s.aE[a], s.kE[a], s.bH[a], s.aH[a], s.kH[a] = [dptr(t) for t in tabs]
        self.psi = [np.zeros(n, np.float64) for _ in range(18)]
        for q in range(18):
            s.psi[q] = dptr(self.psi[q])
        self.s = s

    @property
    def ptr(self):
        return C.byref(self.s)


def step_cpml(grid, cpml, dt):
    lib().orc_step_cpml(grid.ptr, cpml.ptr, dt)


def complex_exp(re, im):
    out = np.zeros(2)
    lib().orc_complex_exp(re, im, dptr(out))
    return complex(out[0], out[1])
