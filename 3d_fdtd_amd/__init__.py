"""Python view of libfdtd3d.so, the MI355X FDTD library with the reference's C surface.

The product is the C-ABI shared library (include/*.h); this module only loads it through
ctypes and mirrors the reference struct layouts (timefield3d.h:15-27, mode.h:13-22,
matrix.h:12-16, complex.h:11-15, adjoint3d.h:16-55) so that tests and bench.py call exactly
what a C driver calls.  There is no fallback: if the library is missing or does not load,
every entry point raises.

Import with ``importlib.import_module("3d_fdtd_amd")`` (the package name starts with a digit).
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libfdtd3d.so")


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


_d, _i, _f = C.c_double, C.c_int, C.c_float
_P = C.POINTER

# name -> (argtypes, restype); the C declarations are in include/*.h
SIGNATURES = {
    # timefield3d.h
    "timefield3d_new": ([_i, _i, _i, _d], TimeField),
    "timefield3d_free": ([TimeField], None),
    "timefield3d_copy": ([TimeField], TimeField),
    "timefield3d_zero_fields": ([TimeField], None),
    "timefield3d_get_fields": ([TimeField], _P(_d)),
    "timefield3d_get_edge_fields": ([TimeField], _P(_d)),
    "timefield3d_update_old_edge_fields": ([TimeField], None),
    "timefield3d_set_background": ([TimeField, _f, _f], None),
    "timefield3d_set_top_half_plane": ([TimeField, _f, _f, _i], None),
    "timefield3d_set_background_permittivity": ([TimeField, _f], None),
    "timefield3d_set_background_permeability": ([TimeField, _f], None),
    "timefield3d_set_background_electric_conductivity": ([TimeField, _f], None),
    "timefield3d_set_background_magnetic_conductivity": ([TimeField, _f], None),
    "timefield3d_set_background_complex_permittivity": ([TimeField, _d, _f, _f], None),
    "timefield3d_set_prism_permittivity": ([TimeField, _f, _i, _i, _i, _i, _i, _i], None),
    "timefield3d_set_prism_electric_conductivity": ([TimeField, _f, _i, _i, _i, _i, _i, _i], None),
    "timefield3d_set_prism_complex_permittivity": ([TimeField, _d, _f, _f, _i, _i, _i, _i, _i, _i], None),
    "timefield3d_set_waveguide": ([TimeField, _f, _i, _i, _i, _i, _i, _i], None),
    "timefield3d_set_taper": ([TimeField, _d, _d, _i, _i, _i, _i, _i, _i, _i, _i], None),
    "timefield3d_apply_H_timestep": ([TimeField, _d], None),
    "timefield3d_apply_E_timestep": ([TimeField, _d], None),
    "timefield3d_apply_H_timestep_zsection": ([TimeField, _d, _i, _i], None),
    "timefield3d_apply_E_timestep_zsection": ([TimeField, _d, _i, _i], None),
    "timefield3d_apply_H_first_order_abc": ([TimeField, _P(_d), _d], None),
    "timefield3d_apply_E_first_order_abc": ([TimeField, _P(_d), _d], None),
    "timefield3d_apply_H_first_order_abc_zsection": ([TimeField, _d, _i, _i], None),
    "timefield3d_apply_E_first_order_abc_zsection": ([TimeField, _d, _i, _i], None),
    "timefield3d_apply_timestep": ([TimeField, _d], None),
    "timefield3d_apply_timestep_omp": ([TimeField, _d], None),
    "timefield3d_apply_timestep_omp_full": ([TimeField, _d], None),
    "timefield3d_update_fourier_fields": ([TimeField, _d], None),
    "timefield3d_update_fourier_fields_y_slice": ([TimeField, _d, _i], None),
    "timefield3d_update_fourier_fields_zsection": ([TimeField, _d, _i, _i], None),
    "timefield3d_update_fourier_fields_y_slice_zsection": ([TimeField, _d, _i, _i, _i], None),
    "timefield3d_update_fourier_fields_omp": ([TimeField, _d], None),
    "timefield3d_update_fourier_fields_y_slice_omp": ([TimeField, _d, _i], None),
    "timefield3d_mode_source": ([TimeField, _i, _i, _i, Mode, _d, _d, _i], None),
    "timefield3d_mode_source_adjoint": ([TimeField, _i, _i, _i, Mode, _d, _i], None),
    "timefield3d_TE_gaussian_source": ([TimeField, _i, _i, _i, _d, _d, _d, _d], None),
    "timefield3d_TM_gaussian_source": ([TimeField, _i, _i, _i, _d, _d, _d, _d], None),
    "timefield3d_mode_overlap": ([TimeField, _i, _i, _i, Mode, _d, _i, _P(Complex)], Complex),
    "timefield3d_file_write_permittivity_y_slice": ([TimeField, C.c_char_p, _i, _i, _i, _i, _i, _d, _d], None),
    **{f"timefield3d_plot_{q}_{a}_slice{suffix}": ([TimeField, _i], None)
       for q, suffix in (("Ex", ""), ("Ey", ""), ("Ez", ""), ("Ex", "_fourier"), ("permittivity", ""))
       for a in "xyz"},
    # mode.h
    "mode_new": ([_i, _i, _d], Mode),
    "mode_free": ([Mode], None),
    "mode_copy": ([Mode], Mode),
    "mode_smul": ([Complex, Mode], None),
    "mode_gaussian": ([_i, _i, _d, Complex, Complex, _d, _d, _i], Mode),
    "mode_power_inner_product": ([Mode, Mode], Complex),
    "mode_power_normalize": ([Mode], None),
    "mode_H_inner_product": ([Mode, Mode], Complex),
    "mode_H_normalize": ([Mode], None),
    "mode_inner_product": ([Mode, Mode], Complex),
    "mode_norm": ([Mode], _d),
    "mode_normalize": ([Mode], None),
    "mode_apply_zero_gradient_bc": ([Mode], None),
    "mode_set_background": ([Mode, Complex, Complex], None),
    "mode_make_rectangle": ([Mode, Complex, Complex, _i, _i, _i, _i], None),
    "mode_set_pml": ([Mode, _d, _i, _i], None),
    "mode_apply_bc": ([Mode, _i], None),
    "mode_apply_beta_squared_operator": ([Mode, _d], Mode),
    "mode_solve_beta_squared": ([_P(Mode), _d, _d, _P(Mode), _i, _i, _i, _i], Complex),
    # matrix.h
    "matrix_new": ([_i, _i], Matrix),
    "matrix_free": ([Matrix], None),
    "matrix_diagonalize": ([Matrix, _d, _i], Matrix),
    "matrix_get_submatrix": ([Matrix, _i, _i], Matrix),
    "matrix_determinant": ([Matrix], Complex),
    "matrix_diagonalize_non_hermitian": ([Matrix, _d, _i], Matrix),
    "matrix_diagonalize_2_by_2": ([Matrix], Matrix),
    "matrix_eigenvector": ([Matrix, Complex], Matrix),
    # complex.h
    "complex_exp": ([Complex], Complex),
    "complex_angle": ([Complex], _d),
    "complex_sqrt": ([Complex], Complex),
    "complex_div": ([Complex, Complex], Complex),
    # adjoint3d.h
    "ajoint3d_new": ([_i, _i, _i, _i, _i, _i, _i], Block),
    "adjoint3d_free": ([Block], None),
    "adjoint3d_store_fwd_fields": ([Block, TimeField, _i], None),
    "adjoint3d_update_gradient": ([Block, TimeField, _i], None),
    "adjoint3d_reset_gradient": ([Block], None),
    "adjoint3d_reset_forward_fields": ([Block], None),
    "adjoint3d_adjust_permittivity": ([Block, TimeField, _d, _d, _d, _d], None),
    "adjoint3d_taper_calculate_gradients": ([Taper, Block, _d], None),
    "adjoint3d_taper_new": ([_d, _d, _i, _i, _i, _P(_i), _P(_i), _i, _i, _i], Taper),
    "adjoint3d_taper_insert": ([TimeField, Taper], None),
    "adjoint3d_taper_store_forward_fields": ([Taper, TimeField, _i], None),
    "adjoint3d_taper_store_adjoint_fields": ([Taper, TimeField, _i], None),
    "adjoint3d_taper_calculate_ep_gradient": ([Taper], None),
    "adjoint3d_taper_calculate_gradients_from_ep_gradient": ([Taper, _d], None),
    "adjoint3d_taper_update_asymmetric": ([Taper, _d, _d, _i], _i),
    "adjoint3d_taper_update_symmetric": ([Taper, _d, _d, _i], _i),
    # fdtd3d_ext.h
    "fdtd3d_ext_sync_to_device": ([TimeField], _i),
    "fdtd3d_ext_sync_from_device": ([TimeField], _i),
    "fdtd3d_ext_upload_source_plane": ([TimeField, _i], _i),
    "fdtd3d_ext_timer_start": ([TimeField], _i),
    "fdtd3d_ext_timer_stop": ([TimeField, _P(_d)], _i),
    "fdtd3d_ext_synchronize": ([TimeField], _i),
    "fdtd3d_ext_steps": ([TimeField, _d, _i], _i),
    "fdtd3d_ext_last_error": ([], C.c_char_p),
    "fdtd3d_ext_kernel_timing": ([TimeField, _i], _i),
    "fdtd3d_ext_kernel_times": ([TimeField, _P(_i), _P(_d)], _i),
    "fdtd3d_ext_set_step_kernel": ([TimeField, _i, _i], _i),
    "fdtd3d_ext_count_nonfinite": ([TimeField, _P(C.c_longlong)], _i),
    "fdtd3d_ext_device_materials": ([TimeField, _P(C.c_float), _P(C.c_float), _P(C.c_float)], _i),
    "fdtd3d_ext_material_counts": ([TimeField, _P(C.c_longlong)], _i),
    "fdtd3d_ext_set_finite_check": ([TimeField, _i], _i),
    "fdtd3d_ext_set_cpml": ([TimeField, _P(_i), _i, _d, _d, _d], _i),
    "fdtd3d_ext_set_mur": ([TimeField], _i),
    "fdtd3d_ext_run_clear": ([TimeField], _i),
    "fdtd3d_ext_run_add_mode_source": ([TimeField, _i, _i, _i, Mode, _i, _i, _i, _P(_d), _P(_d)], _i),
    "fdtd3d_ext_run_add_overlap": ([TimeField, _i, _i, _i, Mode, _i, _i, _P(_d)], _i),
    "fdtd3d_ext_run_add_dft_plane": ([TimeField, _i, _i, _i, _P(_d)], _i),
    "fdtd3d_ext_run": ([TimeField, _d, _i, _i, _d], _i),
    "fdtd3d_ext_run_overlap": ([TimeField, _i, _P(Complex), _P(Complex)], _i),
    "fdtd3d_ext_run_dft": ([TimeField, _i, _P(Complex)], _i),
    "fdtd3d_ext_run_add_dft_box": ([TimeField, _P(_i), _i, _i, _P(_d)], _i),
    "fdtd3d_ext_run_add_adjoint_source": ([TimeField, _i, _i, _i, Mode, _i, _d, _i, _P(_d), _P(_d)], _i),
    "fdtd3d_ext_dft_gradient": ([TimeField, _i, TimeField, _i, _i, Complex, _P(_d)], _i),
    # device-level C-ABI (include/fdtd3d_dev.h): grids without host mirrors, batched runs
    "fdtd_dev_last_error": ([], C.c_char_p),
    "fdtd_dev_create": ([_i, _i, _i, _d], C.c_void_p),
    "fdtd_dev_destroy": ([C.c_void_p], None),
    "fdtd_dev_set_materials": ([C.c_void_p, _P(C.c_float), _P(C.c_float), _P(C.c_float)], _i),
    "fdtd_dev_get_fields": ([C.c_void_p, _P(_d)], _i),
    "fdtd_dev_set_fields": ([C.c_void_p, _P(_d)], _i),
    "fdtd_dev_count_nonfinite": ([C.c_void_p, _P(C.c_longlong)], _i),
    "fdtd_dev_paint_taper": ([C.c_void_p, C.c_float, C.c_float] + [_i] * 8, _i),
    "fdtd_dev_paint_prism": ([C.c_void_p, C.c_float, C.c_float] + [_i] * 7, _i),
    "fdtd_dev_get_materials": ([C.c_void_p, _P(C.c_float), _P(C.c_float), _P(C.c_float)], _i),
    "fdtd_dev_material_counts": ([C.c_void_p, _P(C.c_longlong)], _i),
    "fdtd_dev_copy_planes": ([C.c_void_p, _i, C.c_void_p, _i, C.c_uint], _i),
    "fdtd_dev_set_streaming_loads": ([C.c_void_p, _i], _i),
    "fdtd_dev_set_materials_extruded": ([C.c_void_p, _i, _P(_f), _P(_f), _P(_f), _P(C.c_uint16), _i, _i, _i],
                                        _i),
    "fdtd_dev_update_E_cpml": ([C.c_void_p, _d], _i),
    "fdtd_dev_update_H_cpml": ([C.c_void_p, _d], _i),
    "fdtd_dev_step": ([C.c_void_p, _d, _i], _i),
    "fdtd_dev_set_source_plane": ([C.c_void_p, _i, _P(_f), _P(C.c_ubyte)], _i),
    "fdtd_dev_set_finite_check": ([C.c_void_p, _i], _i),
    "fdtd_dev_cpml_on": ([C.c_void_p, _P(_i), _i, _d, _d, _d], _i),
    "fdtd_dev_set_step_kernel": ([C.c_void_p, _i, _i], _i),
    "fdtd_dev_run_add_source": ([C.c_void_p, _i, _i, _i, _P(_d), _i, _i, _d, _i, _i, _i, _P(_d), _P(_d)], _i),
    "fdtd_dev_run_add_overlap": ([C.c_void_p, _i, _i, _i, _P(_d), _i, _i, _d, _i, _i, _P(_d)], _i),
    "fdtd_dev_run": ([C.c_void_p, _d, _i, _i, _d], _i),
    "fdtd_dev_run_overlap_result": ([C.c_void_p, _i, _P(_d), _P(_d)], _i),
    "fdtd_dev_run_add_dft_box": ([C.c_void_p, _P(_i), _i, _i, _P(_d)], _i),
    "fdtd_dev_run_add_adjoint_source": ([C.c_void_p, _i, _i, _i, _P(_d), _i, _i, _d, _i, _d, _i, _P(_d), _P(_d)],
                                        _i),
    "fdtd_dev_run_dft_result": ([C.c_void_p, _i, _P(_d)], _i),
    "fdtd_dev_dft_gradient": ([C.c_void_p, _i, C.c_void_p, _i, _i, _d, _d, _P(_d)], _i),
    "fdtd_dev_synchronize": ([C.c_void_p], _i),
    "fdtd_dev_mode_overlap": ([C.c_void_p, _i, _i, _i, _P(_d), _i, _i, _d, _i, _P(_d)], _i),
    "fdtd_dev_mode_overlap_noalloc": ([C.c_void_p, _i, _i, _i, C.c_void_p, _i, _i, _d, _i, C.c_void_p], _i),
    "fdtd_dev_release_cached_memory": ([], _i),
    "fdtd_dev_kernel_timing": ([C.c_void_p, _i], _i),
    "fdtd_dev_kernel_times": ([C.c_void_p, _P(_i), _P(_d)], _i),
    "fdtd_dev_timer_start": ([C.c_void_p], _i),
    "fdtd_dev_timer_stop": ([C.c_void_p, _P(_d)], _i),
    "fdtd_dev_batch_new": ([_P(C.c_void_p), _i], C.c_void_p),
    "fdtd_dev_batch_free": ([C.c_void_p], None),
    "fdtd_dev_batch_set_shape": ([C.c_void_p, _i, _i, _i], _i),
    "fdtd_dev_batch_run": ([C.c_void_p, _d, _i, _i], _i),
    "fdtd_dev_batch_info": ([C.c_void_p, _P(_i), _P(C.c_longlong), _P(_i), _P(_i), _P(_d)], _i),
    "fdtd_dev_batch_timer_start": ([C.c_void_p], _i),
    "fdtd_dev_batch_timer_stop": ([C.c_void_p, _P(_d)], _i),
    "fdtd_dev_batch_kernel_timing": ([C.c_void_p, _i], _i),
    "fdtd_dev_batch_kernel_times": ([C.c_void_p, _P(_i), _P(_d)], _i),
    # transmission surrogate (include/fdtd3d_mlp.h)
    "fdtd_mlp_new": ([_i, _i, _i, _i], C.c_void_p),
    "fdtd_mlp_free": ([C.c_void_p], None),
    "fdtd_mlp_param_count": ([C.c_void_p], C.c_longlong),
    "fdtd_mlp_set_params": ([C.c_void_p, _P(C.c_float)], _i),
    "fdtd_mlp_get_params": ([C.c_void_p, _P(C.c_float)], _i),
    "fdtd_mlp_forward": ([C.c_void_p, _P(C.c_float), _i, _P(C.c_float)], _i),
    "fdtd_mlp_loss_grad": ([C.c_void_p, _P(C.c_float), _P(C.c_float), _i, _P(_d), _P(C.c_float)], _i),
    "fdtd_mlp_adam_step": ([C.c_void_p, _P(C.c_float), _P(C.c_float), _i, C.c_float, C.c_float, C.c_float, C.c_float,
                            _P(_d)], _i),
    "fdtd_mlp_train": ([C.c_void_p, _P(C.c_float), _P(C.c_float), _i, _i, C.c_float, C.c_float, C.c_float, C.c_float,
                        _P(_d), _P(_d)], _i),
    "fdtd_mlp_gemm_timing": ([C.c_void_p, _i], _i),
    "fdtd_mlp_gemm_times": ([C.c_void_p, _P(_i), _P(_d), _P(_d)], _i),
    "fdtd_mlp_timer_start": ([C.c_void_p], _i),
    "fdtd_mlp_timer_stop": ([C.c_void_p, _P(_d)], _i),
}

_lib = None


def lib():
    """Load libfdtd3d.so (built by __graft_entry__.build()); raise if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (no CPU fallback exists)")
    L = C.CDLL(LIB_PATH)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(rc, what="fdtd3d call"):
    if rc != 0:
        raise RuntimeError(f"{what}: {lib().fdtd3d_ext_last_error().decode()}")


def _view(ptr, count, ctype, dtype):
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(count,)).view(dtype)


def arrays(tf):
    """numpy views (no copy) of a grid's host arrays, reference layout."""
    n = tf.i * tf.j * tf.k
    return {
        "ep": _view(tf.ep, 3 * n, C.c_float, np.float32),
        "mu": _view(tf.mu, 3 * n, C.c_float, np.float32),
        "sig": _view(tf.sig, 6 * n, C.c_float, np.float32),
        "fields": _view(tf.fields, 6 * n, C.c_double, np.float64),
        "fourier": _view(tf.fourier_fields, 12 * n, C.c_double, np.float64),
        "sources": _view(tf.sources, 6 * n, C.c_float, np.float32),
    }


def mode_fields(md):
    """Interleaved (re, im) view of a mode's 6*i*j complex fields."""
    return _view(md.fields.m, 12 * md.i * md.j, C.c_double, np.float64)


def matrix_view(mat, n):
    """Interleaved (re, im) view of the first n complex entries of a Matrix."""
    return _view(mat.m, 2 * n, C.c_double, np.float64)


def solve_mode(ep, disc, omega, tol=1e-9, size=20, confinement=2, pml=None, lateral_symmetry=0, prev=()):
    """Fundamental mode of a cross-section by the library's mode solver (mode_solve_beta_squared).

    ep: complex permittivity [j][i][3] (x, y, z components); pml: None or (max_ep_im, order,
    thickness) for mode_set_pml; prev: modes (Mode) the result is kept orthogonal to.
    Returns (beta, Mode); the caller frees the Mode (lib().mode_free)."""
    ep = np.asarray(ep, np.complex128)
    mj, mi = ep.shape[:2]
    md = lib().mode_new(mi, mj, disc)
    e = matrix_view(md.ep, 3 * mi * mj).reshape(-1, 2)
    e[:, 0] = ep.reshape(-1).real
    e[:, 1] = ep.reshape(-1).imag
    m = matrix_view(md.mu, 3 * mi * mj).reshape(-1, 2)
    m[:, 0] = 1.0
    m[:, 1] = 0.0
    if pml is not None:
        lib().mode_set_pml(md, float(pml[0]), int(pml[1]), int(pml[2]))
    prev = list(prev)
    arr = (Mode * max(len(prev), 1))(*prev) if prev else None
    beta = lib().mode_solve_beta_squared(C.byref(md), omega, tol, arr, len(prev), size, confinement, lateral_symmetry)
    return complex(beta.re, beta.im), md


def mode_from_array(mi, mj, disc, values):
    """New mode whose fields are the interleaved complex array `values` (12*mi*mj doubles)."""
    md = lib().mode_new(mi, mj, disc)
    mode_fields(md)[:] = values
    return md


class Grid:
    """Owns one timefield3d (host mirrors + device state)."""

    def __init__(self, I, J, K, disc):
        self.tf = lib().timefield3d_new(I, J, K, disc)
        self.I, self.J, self.K, self.disc = I, J, K, disc
        self.a = arrays(self.tf)

    def free(self):
        if self.tf is not None:
            lib().timefield3d_free(self.tf)
            self.tf = None

    def push(self):
        check(lib().fdtd3d_ext_sync_to_device(self.tf), "sync_to_device")

    def pull(self):
        check(lib().fdtd3d_ext_sync_from_device(self.tf), "sync_from_device")
        return self.a

    def step(self, dt, n=1):
        if n == 1:
            lib().timefield3d_apply_timestep_omp_full(self.tf, dt)
        else:
            check(lib().fdtd3d_ext_steps(self.tf, dt, n), "steps")

    def fields(self):
        self.pull()
        return self.a["fields"]


def normalised_mode(mi, mj, disc, values):
    """Power-normalised copy (mode_power_normalize, mode.c semantics) of an interleaved mode array."""
    md = mode_from_array(mi, mj, disc, values)
    lib().mode_power_normalize(md)
    out = mode_fields(md).copy()
    lib().mode_free(md)
    return out


def dev_check(rc, what="fdtd_dev call"):
    if rc != 0:
        raise RuntimeError(f"{what}: {lib().fdtd_dev_last_error().decode()}")


class DevGrid:
    """A device-only grid (fdtd_dev, no host mirrors): the unit of batched runs."""

    def __init__(self, I, J, K, disc, ep, mu=None, sig=None, extruded=None):
        """ep/mu/sig: host AoS materials (set_materials); or extruded = (table [nmat][3] of ep,
        ids [K][I], j0, j1, background id) for a geometry rasterised on the device."""
        n = I * J * K
        self.I, self.J, self.K, self.disc = I, J, K, disc
        self.d = lib().fdtd_dev_create(I, J, K, disc)
        if not self.d:
            raise RuntimeError("fdtd_dev_create: " + lib().fdtd_dev_last_error().decode())
        fp = C.POINTER(C.c_float)

        def arr(a, size):  # None -> NULL (the library's mu = 1 / sig = 0 defaults)
            if a is None:
                return None, None
            a = np.ascontiguousarray(a, np.float32)
            if a.size != size:
                raise ValueError(f"material array of {a.size} values for {size}")
            return a, a.ctypes.data_as(fp)

        try:
            if extruded is not None:  # (table_ep [nmat][3], ids [K][I], j0, j1, background)
                table, ids, j0, j1, bg = extruded
                table = np.ascontiguousarray(table, np.float32)
                ids = np.ascontiguousarray(ids, np.uint16)
                if ids.size != K * I:
                    raise ValueError(f"{ids.size} extruded ids for {K * I}")
                dev_check(lib().fdtd_dev_set_materials_extruded(
                    self.d, table.shape[0], table.ctypes.data_as(fp), None, None,
                    ids.ctypes.data_as(C.POINTER(C.c_uint16)), j0, j1, bg), "set_materials_extruded")
                return
            if ep is None:
                raise ValueError("DevGrid needs a permittivity array")
            ep, ep_p = arr(ep, 3 * n)
            mu, mu_p = arr(mu, 3 * n)
            sig, sig_p = arr(sig, 6 * n)
            dev_check(lib().fdtd_dev_set_materials(self.d, ep_p, mu_p, sig_p), "set_materials")
        except Exception:
            self.free()
            raise

    def cpml(self, npml, order=3, R0=1e-8, kappa_max=1.0, alpha_max=0.05):
        dev_check(lib().fdtd_dev_cpml_on(self.d, (C.c_int * 6)(*npml), order, R0, kappa_max, alpha_max), "cpml_on")

    def add_source(self, cx, cy, cz, mode, mi, mj, mdisc, amplitude, phase, reverse=0, adjoint=0):
        dp = C.POINTER(C.c_double)
        amplitude = np.ascontiguousarray(amplitude, np.float64)
        phase = np.ascontiguousarray(phase, np.float64)
        mode = np.ascontiguousarray(mode, np.float64)
        sid = lib().fdtd_dev_run_add_source(self.d, cx, cy, cz, mode.ctypes.data_as(dp), mi, mj, mdisc, reverse, adjoint,
                                            len(amplitude), amplitude.ctypes.data_as(dp), phase.ctypes.data_as(dp))
        dev_check(0 if sid >= 0 else sid, "run_add_source")
        return sid

    def add_overlap(self, cx, cy, cz, mode, mi, mj, mdisc, omegas, reverse=0):
        dp = C.POINTER(C.c_double)
        omegas = np.ascontiguousarray(omegas, np.float64)
        mode = np.ascontiguousarray(mode, np.float64)
        mid = lib().fdtd_dev_run_add_overlap(self.d, cx, cy, cz, mode.ctypes.data_as(dp), mi, mj, mdisc, reverse,
                                             len(omegas), omegas.ctypes.data_as(dp))
        dev_check(0 if mid >= 0 else mid, "run_add_overlap")
        return mid

    def add_dft_box(self, box, comp_mask, omegas):
        dp = C.POINTER(C.c_double)
        omegas = np.ascontiguousarray(omegas, np.float64)
        bid = lib().fdtd_dev_run_add_dft_box(self.d, (C.c_int * 6)(*box), comp_mask, len(omegas),
                                             omegas.ctypes.data_as(dp))
        dev_check(0 if bid >= 0 else bid, "run_add_dft_box")
        return bid

    def add_adjoint_source(self, cx, cy, cz, mode, mi, mj, mdisc, omega_dt, amplitude, phase, reverse=0):
        dp = C.POINTER(C.c_double)
        amplitude = np.ascontiguousarray(amplitude, np.float64)
        phase = np.ascontiguousarray(phase, np.float64)
        mode = np.ascontiguousarray(mode, np.float64)
        sid = lib().fdtd_dev_run_add_adjoint_source(self.d, cx, cy, cz, mode.ctypes.data_as(dp), mi, mj, mdisc, reverse,
                                                    omega_dt, len(amplitude), amplitude.ctypes.data_as(dp),
                                                    phase.ctypes.data_as(dp))
        dev_check(0 if sid >= 0 else sid, "run_add_adjoint_source")
        return sid

    def gradient(self, bid, adjoint, aid, freq, weight, nb):
        """Re[weight * sum_m F_m A_m] over the box (this grid forward, `adjoint` the adjoint grid)."""
        out = np.zeros(nb)
        dev_check(lib().fdtd_dev_dft_gradient(self.d, bid, adjoint.d, aid, freq, weight.real, weight.imag,
                                              out.ctypes.data_as(C.POINTER(C.c_double))), "dft_gradient")
        return out

    def run(self, dt, n0, n, t_offset=0.0):
        dev_check(lib().fdtd_dev_run(self.d, dt, n0, n, t_offset), "run")

    def kernel_timing(self, every):
        """Bracket every `every`-th split E / H launch with dispatch-recorded events (0: off)."""
        dev_check(lib().fdtd_dev_kernel_timing(self.d, every), "kernel_timing")

    def kernel_times(self):
        """Mean durations (us) of the timed E and H launches and their counts; clears the record."""
        cnt, ms = (C.c_int * 3)(), (C.c_double * 3)()
        dev_check(lib().fdtd_dev_kernel_times(self.d, cnt, ms), "kernel_times")
        return [ms[q] / cnt[q] * 1e3 if cnt[q] else 0.0 for q in range(2)], [cnt[0], cnt[1]]

    def timer_start(self):
        dev_check(lib().fdtd_dev_timer_start(self.d), "timer_start")

    def timer_stop(self):
        ms = C.c_double()
        dev_check(lib().fdtd_dev_timer_stop(self.d, C.byref(ms)), "timer_stop")
        return ms.value

    def overlap(self, mid, nfreq):
        cum = np.zeros(2 * nfreq)
        last = np.zeros(2)
        dp = C.POINTER(C.c_double)
        dev_check(lib().fdtd_dev_run_overlap_result(self.d, mid, cum.ctypes.data_as(dp), last.ctypes.data_as(dp)),
                  "overlap result")
        return cum[0::2] + 1j * cum[1::2], complex(last[0], last[1])

    def dft_box(self, bid, ncells, nfreq=1):
        """The box's FP64 DFT accumulators as interleaved (re, im) doubles (fdtd_dev_run_dft_result)."""
        out = np.zeros(2 * 3 * ncells * nfreq)
        dev_check(lib().fdtd_dev_run_dft_result(self.d, bid, out.ctypes.data_as(C.POINTER(C.c_double))), "dft result")
        return out

    def fields(self):
        out = np.zeros(6 * self.I * self.J * self.K)
        dev_check(lib().fdtd_dev_get_fields(self.d, out.ctypes.data_as(C.POINTER(C.c_double))), "get_fields")
        return out

    def set_fields(self, f):
        """Fields [k][j][i][6] (Ex, Ey, Ez, Hx, Hy, Hz), stored as FP32 on the device."""
        f = np.ascontiguousarray(f, np.float64)
        if f.size != 6 * self.I * self.J * self.K:
            raise ValueError(f"{f.size} field values for {6 * self.I * self.J * self.K}")
        dev_check(lib().fdtd_dev_set_fields(self.d, f.ctypes.data_as(C.POINTER(C.c_double))), "set_fields")

    def free(self):
        if self.d:
            lib().fdtd_dev_destroy(self.d)
            self.d = None


class Batch:
    """Lockstep batched run over DevGrids (fdtd_dev_batch_*)."""

    def __init__(self, grids):
        self.grids = list(grids)
        arr = (C.c_void_p * len(self.grids))(*[g.d for g in self.grids])
        self.b = lib().fdtd_dev_batch_new(arr, len(self.grids))
        if not self.b:
            raise RuntimeError("fdtd_dev_batch_new: " + lib().fdtd_dev_last_error().decode())

    def run(self, dt, n0, n):
        dev_check(lib().fdtd_dev_batch_run(self.b, dt, n0, n), "batch_run")

    def set_shape(self, kchunk, unroll=True, persistent=True):
        dev_check(lib().fdtd_dev_batch_set_shape(self.b, kchunk, int(unroll), int(persistent)), "batch_set_shape")

    @staticmethod
    def shape_name(kchunk, unroll, persistent):
        return f"kc{kchunk}-{'unrolled' if unroll else 'loop'}-{'persistent' if persistent else 'full'}"

    def info(self):
        sh, nt = (C.c_int * 3)(), C.c_int()
        cells = C.c_longlong()
        tk, tm = (C.c_int * 16)(), (C.c_double * 16)()
        dev_check(lib().fdtd_dev_batch_info(self.b, sh, C.byref(cells), C.byref(nt), tk, tm), "batch_info")
        tuned = {self.shape_name(tk[q] % 16, tk[q] & 16, tk[q] & 32): float(tm[q]) for q in range(nt.value)}
        return {"shape": self.shape_name(sh[0], sh[1], sh[2]), "kchunk": sh[0], "cells": cells.value, "tuned": tuned}

    def free(self):
        if self.b:
            lib().fdtd_dev_batch_free(self.b)
            self.b = None


class Mlp:
    """The transmission surrogate (fdtd_mlp_*): BF16-MFMA MLP trained and run on the GPU."""

    def __init__(self, n_in, width, depth, n_out, params=None):
        self.dims = (n_in, width, depth, n_out)
        self.m = lib().fdtd_mlp_new(n_in, width, depth, n_out)
        if not self.m:
            raise RuntimeError("fdtd_mlp_new: " + lib().fdtd_dev_last_error().decode())
        self.n = int(lib().fdtd_mlp_param_count(self.m))
        if params is not None:
            self.set_params(params)

    @staticmethod
    def _fp(a):
        return a.ctypes.data_as(C.POINTER(C.c_float))

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n
        dev_check(lib().fdtd_mlp_set_params(self.m, self._fp(p)), "mlp set_params")

    def params(self):
        p = np.zeros(self.n, np.float32)
        dev_check(lib().fdtd_mlp_get_params(self.m, self._fp(p)), "mlp get_params")
        return p

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], self.dims[3]), np.float32)
        dev_check(lib().fdtd_mlp_forward(self.m, self._fp(x), x.shape[0], self._fp(y)), "mlp forward")
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.n, np.float32)
        loss = C.c_double()
        dev_check(lib().fdtd_mlp_loss_grad(self.m, self._fp(x), self._fp(t), x.shape[0], C.byref(loss), self._fp(g)),
                  "mlp loss_grad")
        return loss.value, g

    def adam_step(self, x, t, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = C.c_double()
        dev_check(lib().fdtd_mlp_adam_step(self.m, self._fp(x), self._fp(t), x.shape[0], lr, beta1, beta2, eps,
                                           C.byref(loss)), "mlp adam_step")
        return loss.value

    def train(self, x, t, steps, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        """`steps` full-batch Adam steps on device-resident (x, t); returns (loss before, loss after)."""
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        first, last = C.c_double(), C.c_double()
        dev_check(lib().fdtd_mlp_train(self.m, self._fp(x), self._fp(t), x.shape[0], steps, lr, beta1, beta2, eps,
                                       C.byref(first), C.byref(last)), "mlp train")
        return first.value, last.value

    def gemm_timing(self, on=True):
        dev_check(lib().fdtd_mlp_gemm_timing(self.m, int(on)), "mlp gemm_timing")

    def gemm_times(self):
        """(launches, summed ms, flop) of the GEMM launches timed since gemm_timing(True)."""
        n, ms, fl = C.c_int(), C.c_double(), C.c_double()
        dev_check(lib().fdtd_mlp_gemm_times(self.m, C.byref(n), C.byref(ms), C.byref(fl)), "mlp gemm_times")
        return n.value, ms.value, fl.value

    def timer_start(self):
        dev_check(lib().fdtd_mlp_timer_start(self.m), "mlp timer")

    def timer_stop(self):
        ms = C.c_double()
        dev_check(lib().fdtd_mlp_timer_stop(self.m, C.byref(ms)), "mlp timer")
        return ms.value

    def free(self):
        if self.m:
            lib().fdtd_mlp_free(self.m)
            self.m = None
