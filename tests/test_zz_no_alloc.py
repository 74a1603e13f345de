"""GPU test of the allocation-free overlap entry point (fdtd_dev_mode_overlap_noalloc): through the
C-ABI, against fdtd_dev_mode_overlap (the same reduction kernel with library-allocated scratch) on
the same fields and mode: bit-identical (re, im), on two calls in a row, the second on a smaller
grid and mode window (and the reversed-mode sign), so the shared per-GPU scratch is reused with
fewer block partials.  A plane beyond the scratch limit is refused on the host before any launch.

The test's own device buffers (mode, output) come from the HIP runtime the library is linked
against, through ctypes."""
import ctypes as C
import importlib
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

fd = importlib.import_module("3d_fdtd_amd")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipFree.argtypes = [C.c_void_p]
    return _hip


class DevBuf:
    def __init__(self, nbytes):
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), nbytes) == 0
        self.n = nbytes

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.n
        assert hip().hipMemcpy(self.p, a.ctypes.data, a.nbytes, 1) == 0

    def download(self, a):
        assert hip().hipMemcpy(a.ctypes.data, self.p, a.nbytes, 2) == 0
        return a

    def free(self):
        hip().hipFree(self.p)


def case(I, J, K, mi, mj, cz, reverse, seed):
    rng = np.random.default_rng(seed)
    n = I * J * K
    g = fd.DevGrid(I, J, K, 0.02, rng.uniform(1, 12, 3 * n).astype(np.float32))
    g.set_fields(rng.uniform(-1, 1, 6 * n))
    mode = rng.normal(size=12 * mi * mj)
    return g, mode, (I // 2, J // 2, cz, mi, mj, 0.02, reverse)


def test_noalloc_overlap_equals_allocating_overlap_twice():
    L = fd.lib()
    mode_buf = DevBuf(12 * 30 * 24 * 8)
    out_buf = DevBuf(16)
    try:
        for call, (dims, seed) in enumerate((((40, 32, 30, 30, 24, 17, 0), 1), ((24, 20, 16, 10, 8, 5, 1), 2))):
            I, J, K, mi, mj, cz, rev = dims
            say(f"no-alloc overlap: call {call + 1}, grid {I}x{J}x{K}, mode {mi}x{mj}, plane {cz}, reverse {rev}")
            g, mode, (cx, cy, cz, mi, mj, mdisc, rev) = case(I, J, K, mi, mj, cz, rev, seed)
            try:
                ref = np.zeros(2)
                dp = C.POINTER(C.c_double)
                fd.dev_check(L.fdtd_dev_mode_overlap(g.d, cx, cy, cz, mode.ctypes.data_as(dp), mi, mj, mdisc, rev,
                                                     ref.ctypes.data_as(dp)), "mode_overlap")
                mode_buf.upload(mode)
                out_buf.upload(np.full(2, np.nan))
                fd.dev_check(L.fdtd_dev_mode_overlap_noalloc(g.d, cx, cy, cz, mode_buf.p, mi, mj, mdisc, rev,
                                                             out_buf.p), "mode_overlap_noalloc")
                fd.dev_check(L.fdtd_dev_synchronize(g.d), "synchronize")
                got = out_buf.download(np.zeros(2))
            finally:
                g.free()
            say(f"  allocating {ref}, no-alloc {got}")
            assert np.all(np.isfinite(ref)) and np.abs(ref).max() > 0
            assert np.array_equal(got, ref), (got, ref)
    finally:
        mode_buf.free()
        out_buf.free()


def test_noalloc_overlap_refuses_planes_beyond_its_scratch():
    L = fd.lib()
    g = fd.DevGrid(1100, 1000, 3, 0.02, None, extruded=(np.ones((1, 3), np.float32),
                                                        np.zeros((3, 1100), np.uint16), 0, 0, 0))
    try:
        rc = L.fdtd_dev_mode_overlap_noalloc(g.d, 550, 500, 1, C.c_void_p(16), 4, 4, 0.02, 0, C.c_void_p(16))
        msg = L.fdtd_dev_last_error().decode()
        say("no-alloc overlap on a 1.1 M-cell plane:", rc, msg)
        assert rc == -2 and "limited" in msg
    finally:
        g.free()
