"""ctypes binding of the compiled reference (oracle/_ref/libref3dfdtd.so) -- TEST INFRASTRUCTURE.

The library is the reference C compiled from its own sources by ``make -C oracle ref``
(recipe in oracle/Makefile).  It is used only as a checker: by tests/golden/make_golden.py to
write fixtures, by tests that compare the clean-room oracle with it when it is present, and
by the cpu_baseline leg of bench.py.  Struct layouts follow the reference headers
(timefield3d.h:15-27, mode.h:13-22, matrix.h:12-16, complex.h:11-15).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_LIB = os.path.join(HERE, "_ref", "libref3dfdtd.so")


class Complex(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class Matrix(C.Structure):
    _fields_ = [("i", C.c_int), ("j", C.c_int), ("m", C.POINTER(Complex))]


class Mode(C.Structure):
    _fields_ = [("i", C.c_int), ("j", C.c_int), ("disc", C.c_double), ("pml", Matrix), ("ep", Matrix),
                ("mu", Matrix), ("fields", Matrix)]


class TimeField(C.Structure):
    _fields_ = [("i", C.c_int), ("j", C.c_int), ("k", C.c_int), ("disc", C.c_double),
                ("ep", C.POINTER(C.c_float)), ("mu", C.POINTER(C.c_float)), ("sig", C.POINTER(C.c_float)),
                ("fields", C.POINTER(C.c_double)), ("old_fields", C.POINTER(C.c_double)),
                ("fourier_fields", C.POINTER(Complex)), ("sources", C.POINTER(C.c_float))]


class Taper(C.Structure):
    _fields_ = [("ep", C.c_double), ("mu", C.c_double), ("seg_num", C.c_int), ("start", C.c_int),
                ("end", C.c_int), ("bottom", C.c_int), ("top", C.c_int), ("duration", C.c_int),
                ("left", C.POINTER(C.c_int)), ("right", C.POINTER(C.c_int)),
                ("forward_fields", C.POINTER(C.c_double)), ("adjoint_fields", C.POINTER(C.c_double)),
                ("ep_gradient", C.POINTER(C.c_double)), ("left_grad", C.POINTER(C.c_double)),
                ("right_grad", C.POINTER(C.c_double))]


class Block(C.Structure):
    _fields_ = [("t", C.c_int), ("x_start", C.c_int), ("x_end", C.c_int), ("y_start", C.c_int),
                ("y_end", C.c_int), ("z_start", C.c_int), ("z_end", C.c_int),
                ("forward_fields", C.POINTER(C.c_double)), ("gradient", C.POINTER(C.c_double))]


def available():
    return os.path.exists(REF_LIB)


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not available():
        raise FileNotFoundError(f"{REF_LIB} not built (make -C oracle ref, needs /root/reference)")
    L = C.CDLL(REF_LIB)
    d, i, f = C.c_double, C.c_int, C.c_float
    sig = {
        "timefield3d_new": ([i, i, i, d], TimeField),
        "timefield3d_free": ([TimeField], None),
        "timefield3d_zero_fields": ([TimeField], None),
        "timefield3d_apply_timestep_omp_full": ([TimeField, d], None),
        "timefield3d_apply_timestep": ([TimeField, d], None),
        "timefield3d_update_fourier_fields": ([TimeField, d], None),
        "timefield3d_update_fourier_fields_y_slice": ([TimeField, d, i], None),
        "timefield3d_update_fourier_fields_y_slice_omp": ([TimeField, d, i], None),
        "timefield3d_mode_source": ([TimeField, i, i, i, Mode, d, d, i], None),
        "timefield3d_mode_source_adjoint": ([TimeField, i, i, i, Mode, d, i], None),
        "timefield3d_mode_overlap": ([TimeField, i, i, i, Mode, d, i, C.POINTER(Complex)], Complex),
        "timefield3d_set_background": ([TimeField, f, f], None),
        "timefield3d_set_top_half_plane": ([TimeField, f, f, i], None),
        "timefield3d_set_taper": ([TimeField, d, d, i, i, i, i, i, i, i, i], None),
        "timefield3d_set_waveguide": ([TimeField, f, i, i, i, i, i, i], None),
        "timefield3d_set_prism_permittivity": ([TimeField, f, i, i, i, i, i, i], None),
        "timefield3d_file_write_permittivity_y_slice": ([TimeField, C.c_char_p, i, i, i, i, i, d, d], None),
        "matrix_new": ([i, i], Matrix),
        "matrix_free": ([Matrix], None),
        "matrix_get_submatrix": ([Matrix, i, i], Matrix),
        "matrix_determinant": ([Matrix], Complex),
        "matrix_diagonalize_non_hermitian": ([Matrix, d, i], Matrix),
        "matrix_diagonalize_2_by_2": ([Matrix], Matrix),
        "mode_new": ([i, i, d], Mode),
        "mode_file_print": ([C.c_void_p, Mode, i], None),
        "mode_file_load": ([C.c_void_p], Mode),
        "timefield3d_TE_gaussian_source": ([TimeField, i, i, i, d, d, d, d], None),
        "timefield3d_TM_gaussian_source": ([TimeField, i, i, i, d, d, d, d], None),
        "mode_inner_product": ([Mode, Mode], Complex),
        "mode_norm": ([Mode], d),
        "mode_normalize": ([Mode], None),
        "mode_apply_zero_gradient_bc": ([Mode], None),
        "mode_free": ([Mode], None),
        "mode_gaussian": ([i, i, d, Complex, Complex, d, d, i], Mode),
        "mode_power_normalize": ([Mode], None),
        "mode_power_inner_product": ([Mode, Mode], Complex),
        "mode_smul": ([Complex, Mode], None),
        "mode_set_pml": ([Mode, d, i, i], None),
        "mode_apply_beta_squared_operator": ([Mode, d], Mode),
        "complex_exp": ([Complex], Complex),
        "complex_angle": ([Complex], d),
        "complex_sqrt": ([Complex], Complex),
        "ajoint3d_new": ([i, i, i, i, i, i, i], Block),
        "adjoint3d_store_fwd_fields": ([Block, TimeField, i], None),
        "adjoint3d_update_gradient": ([Block, TimeField, i], None),
        "adjoint3d_adjust_permittivity": ([Block, TimeField, d, d, d, d], None),
        "adjoint3d_taper_calculate_gradients": ([Taper, Block, d], None),
        "adjoint3d_taper_new": ([d, d, i, i, i, C.POINTER(i), C.POINTER(i), i, i, i], Taper),
        "adjoint3d_taper_insert": ([TimeField, Taper], None),
        "adjoint3d_taper_store_forward_fields": ([Taper, TimeField, i], None),
        "adjoint3d_taper_store_adjoint_fields": ([Taper, TimeField, i], None),
        "adjoint3d_taper_calculate_ep_gradient": ([Taper], None),
        "adjoint3d_taper_calculate_gradients_from_ep_gradient": ([Taper, d], None),
        "adjoint3d_taper_update_asymmetric": ([Taper, d, d, i], i),
        "adjoint3d_taper_update_symmetric": ([Taper, d, d, i], i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def n_cells(tf):
    return tf.i * tf.j * tf.k


def view(ptr, count, dtype):
    """numpy view (no copy) of `count` elements behind a ctypes pointer."""
    return np.ctypeslib.as_array(ptr, shape=(count,)).view(dtype)


def tf_arrays(tf):
    n = n_cells(tf)
    return {
        "ep": view(tf.ep, 3 * n, np.float32),
        "mu": view(tf.mu, 3 * n, np.float32),
        "sig": view(tf.sig, 6 * n, np.float32),
        "fields": view(tf.fields, 6 * n, np.float64),
        "old": view(tf.old_fields, 6 * n, np.float64),
        "fourier": np.ctypeslib.as_array(C.cast(tf.fourier_fields, C.POINTER(C.c_double)), shape=(12 * n,)),
        "sources": view(tf.sources, 6 * n, np.float32),
    }


def mode_fields(md):
    """Interleaved (re, im) float64 view of a mode's 6*i*j complex field array."""
    n = 6 * md.i * md.j
    return np.ctypeslib.as_array(C.cast(md.fields.m, C.POINTER(C.c_double)), shape=(2 * n,))


def matrix_array(M):
    """Copy of a matrix's entries as a complex (i, j) numpy array."""
    raw = np.ctypeslib.as_array(C.cast(M.m, C.POINTER(C.c_double)), shape=(2 * M.i * M.j,)) if M.i * M.j else np.zeros(0)
    return (raw[0::2] + 1j * raw[1::2]).reshape(M.i, M.j)


def matrix_from(L, A):
    """New reference matrix holding the complex 2-D array A."""
    A = np.asarray(A, complex)
    M = L.matrix_new(A.shape[0], A.shape[1])
    raw = np.ctypeslib.as_array(C.cast(M.m, C.POINTER(C.c_double)), shape=(2 * A.size,))
    raw[0::2] = A.reshape(-1).real
    raw[1::2] = A.reshape(-1).imag
    return M
