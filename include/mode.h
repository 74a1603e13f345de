/*
 * mode.h -- waveguide mode profiles of the drop-in C surface.
 *
 * Replaces the reference interface mode.h:13-121 for the functions listed here (same struct
 * layout, names, signatures).  A mode is an i x j transverse grid with disc spacing; fields.m
 * holds 6*i*j complex values ordered Ex,Ey,Ez,Hx,Hy,Hz per cell, cell index i + j*mode.i.
 * The FDTD hot path consumes modes through timefield3d_mode_source/_adjoint/_overlap, which
 * power-normalise the caller's mode in place on every call (mode.c:1367).
 *
 * The finite-difference mode solver (mode_solve_beta_squared with its operator, boundary
 * conditions, PML profile and lateral symmetry) is provided; the first-order beta solver, the
 * one-sided PML setters, radial/angular variants and the susceptibility integrals
 * (mode.c:1792-2198, 2399-2795, 3146-3349, 3404-3460, 3618-3788, 3827-4493) are not.
 */
#ifndef _mode_h
#define _mode_h

#include "complex.h"
#include "matrix.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

/* mode.h:13-22 */
struct mode_ {
    int i, j;
    double disc;
    matrix pml;    /* 2*i*j stretch factors sx, sy */
    matrix ep;     /* 3*i*j */
    matrix mu;     /* 3*i*j */
    matrix fields; /* 6*i*j */
};
typedef struct mode_ mode;

mode mode_new(int i, int j, double disc);                                  /* mode.c:4    */
void mode_free(mode psi);                                                  /* mode.c:37   */
mode mode_copy(mode psi);                                                  /* mode.c:45   */
void mode_copy_fields(mode psi, mode phi);                                 /* mode.c:256  */
void mode_smul(complex scalar, mode psi);                                  /* mode.c:78   */
mode mode_linear_combination(mode psi, mode phi, complex scalar);          /* mode.c:96   */
mode mode_linear_combination_3(mode psi, mode phi, mode alpha, complex scalar_phi,
                               complex scalar_alpha);                      /* mode.c:135  */
mode mode_gaussian(int i, int j, double disc, complex ep, complex mu, double omega, double sigma,
                   int polarization);                                      /* mode.c:174  */
void mode_chop_bottom_half(mode psi);                                      /* mode.c:236  */

/* Plots: the reference pipes to gnuplot (mode.c:273-842).  Here they write gnuplot-ready text
 * files into $FDTD3D_PLOT_DIR when that variable is set, and do nothing otherwise. */
void mode_plot(mode psi);
void mode_plot_real(mode psi);
void mode_plot_ex(mode psi);
void mode_plot_hy(mode psi);

void mode_file_print(FILE *mode_data, mode psi, int dimension);            /* mode.c:843  */
mode mode_file_load(FILE *mode_data);                                      /* mode.c:949  */

double mode_max_index(mode psi);                                           /* mode.c:1131 */
void mode_initialize_ones(mode psi);                                       /* mode.c:1150 */

complex mode_inner_product(mode psi, mode phi);                            /* mode.c:1208 */
complex mode_H_inner_product(mode psi, mode phi);                          /* mode.c:1237 */
complex mode_power_inner_product(mode psi, mode phi);                      /* mode.c:1267 */
double mode_norm(mode psi);                                                /* mode.c:1312 */
double mode_H_norm(mode psi);                                              /* mode.c:1321 */
double mode_power_norm(mode psi);                                          /* mode.c:1330 */
void mode_normalize(mode psi);                                             /* mode.c:1339 */
void mode_H_normalize(mode psi);                                           /* mode.c:1353 */
void mode_power_normalize(mode psi);                                       /* mode.c:1367 */

void mode_set_background(mode psi, complex ep, complex mu);                /* mode.c:2199 */
void mode_set_bottom_half_plane(mode psi, complex ep, complex mu, int top);
void mode_set_top_half_plane(mode psi, complex ep, complex mu, int bottom);
void mode_set_left_half_plane(mode psi, complex ep, complex mu, int right);
void mode_set_right_half_plane(mode psi, complex ep, complex mu, int left);
void mode_make_rectangle(mode psi, complex ep, complex mu, int left, int right, int bottom, int top);

void mode_apply_H_curl(mode psi, complex beta, double omega);              /* mode.c:2330 */
void mode_apply_H_div(mode psi, complex beta);                             /* mode.c:3350 */

/* Mode solver: the fundamental (largest Re beta^2) mode of the cross-section in psi (ep, mu, pml
 * set by the caller), orthogonal in the H inner product to prev_modes[0..prev_num); size is the
 * Krylov dimension per restart, tol the change of |n_eff| between restarts that ends the
 * iteration; confinement 1 = slab (zero top/bottom, zero-gradient sides), 2 = zero on every edge.
 * Returns beta; psi holds all six components, H-normalised.  Starts from rand() values.  As in
 * the reference only the restart vector is orthogonalised against prev_modes, so the operator
 * (not normal in the H inner product) regrows their components inside each Krylov space: higher
 * modes found this way are not reliable. */
complex mode_solve_beta_squared(mode *psi, double omega, double tol, mode *prev_modes, int prev_num, int size,
                                int confinement, int lateral_symmetry);        /* mode.c:3461 */
mode mode_apply_beta_squared_operator(mode psi, double omega);             /* mode.c:2796 */
void mode_initialize_general(mode psi);                                    /* mode.c:1164 */
void mode_initialize_rand(mode psi);                                       /* mode.c:1192 */
void mode_apply_zero_bc(mode psi);                                         /* mode.c:1381 */
void mode_apply_zero_bc_vertical(mode psi);                                /* mode.c:1429 */
void mode_apply_zero_bc_horizontal(mode psi);                              /* mode.c:1457 */
void mode_apply_zero_gradient_bc_vertical(mode psi);                       /* mode.c:1532 */
void mode_apply_zero_gradient_bc_horizontal(mode psi);                     /* mode.c:1560 */
/* both pairs of edges; the left edge copies with the reference's one-slot shift (mode.c:1484) */
void mode_apply_zero_gradient_bc(mode psi);
void mode_apply_zero_bc_pml(mode psi);                                     /* mode.c:1587 */
void mode_apply_bc(mode psi, int confinement);                             /* mode.c:1614 */
void mode_set_pml(mode psi, double max_ep_im, int order, int pml_thickness); /* mode.c:1634 */
void mode_apply_lateral_symmetry(mode psi);                                /* mode.c:4494 */
complex mode_true_beta(complex beta);                                      /* mode.c:3789 */
mode mode_true_fields(mode psi);                                           /* mode.c:3798 */

#ifdef __cplusplus
}
#endif
#endif
