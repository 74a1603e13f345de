/*
 * timefield3d.h -- the FDTD engine's drop-in C surface (MI355X implementation).
 *
 * Replaces the reference interface timefield3d.h:15-104 one for one: the struct layout, the
 * function names, their argument meaning and their by-value calling convention are kept so that
 * drivers written against the reference (3d_fdtd.c) compile and link unchanged.
 *
 * What differs underneath: every grid created by timefield3d_new also owns device-resident
 * state on the current HIP device (FP32 structure-of-arrays fields, material-id cells, source
 * planes, monitors).  The time-stepping, source, DFT, overlap and edge-store calls run as
 * hand-written HIP kernels on that state.  The struct's host arrays stay allocated with the
 * reference layout and act as mirrors:
 *   - ep / mu / sig are host-authoritative: the setters below write them and mark the device
 *     copy stale; it is rebuilt before the next device call.  Code that writes these arrays
 *     directly must call fdtd3d_ext_sync_to_device() (fdtd3d_ext.h) afterwards.
 *   - fields / sources / fourier_fields are device-authoritative: they are refreshed by the
 *     calls that read them on the host (get_fields, plots, file output) and by
 *     fdtd3d_ext_sync_from_device().  old_fields is scratch of the step and is not mirrored.
 * Copies of the struct alias the same arrays and therefore the same device state (state is
 * keyed by the fields pointer).  Errors from the HIP runtime are reported on stderr and abort,
 * matching the reference's abort-on-mismatch convention (mode.c:101-105).
 */
#ifndef _timefield3d_h
#define _timefield3d_h

#include "complex.h"
#include "matrix.h"
#include "mode.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

/* timefield3d.h:15-27.  Cell c = i + I*(j + J*k); fields[6c+m] = Ex,Ey,Ez,Hx,Hy,Hz;
 * ep/mu[3c+m]; sig/sources[6c+m] (electric m<3, magnetic m>=3). */
struct timefield3d_ {
    int i, j, k;
    double disc;
    float *ep;
    float *mu;
    float *sig;
    double *fields;
    double *old_fields;
    complex *fourier_fields;
    float *sources;
};
typedef struct timefield3d_ timefield3d;

/* lifecycle (timefield3d.c:6-346) */
timefield3d timefield3d_new(int i, int j, int k, double disc);
void timefield3d_free(timefield3d psi);
timefield3d timefield3d_copy(timefield3d psi);
void timefield3d_zero_fields(timefield3d psi);
double *timefield3d_get_fields(timefield3d psi);
double *timefield3d_get_edge_fields(timefield3d psi);
void timefield3d_update_old_edge_fields(timefield3d psi);

/* plots (timefield3d.c:348-1354): the reference pipes to gnuplot; here they write a text slice
 * into $FDTD3D_PLOT_DIR when that variable is set, and otherwise do nothing. */
void timefield3d_plot_Ex_x_slice(timefield3d psi, int slice_index);
void timefield3d_plot_Ex_y_slice(timefield3d psi, int slice_index);
void timefield3d_plot_Ex_z_slice(timefield3d psi, int slice_index);
void timefield3d_plot_Ey_x_slice(timefield3d psi, int slice_index);
void timefield3d_plot_Ey_y_slice(timefield3d psi, int slice_index);
void timefield3d_plot_Ey_z_slice(timefield3d psi, int slice_index);
void timefield3d_plot_Ez_x_slice(timefield3d psi, int slice_index);
void timefield3d_plot_Ez_y_slice(timefield3d psi, int slice_index);
void timefield3d_plot_Ez_z_slice(timefield3d psi, int slice_index);
void timefield3d_plot_Ex_x_slice_fourier(timefield3d psi, int slice_index);
void timefield3d_plot_Ex_y_slice_fourier(timefield3d psi, int slice_index);
void timefield3d_plot_Ex_z_slice_fourier(timefield3d psi, int slice_index);
void timefield3d_plot_permittivity_x_slice(timefield3d psi, int slice_index);
void timefield3d_plot_permittivity_y_slice(timefield3d psi, int slice_index);
void timefield3d_plot_permittivity_z_slice(timefield3d psi, int slice_index);

/* material setters, host side (timefield3d.c:1356-1535) */
void timefield3d_set_background(timefield3d psi, float ep, float mu);
void timefield3d_set_top_half_plane(timefield3d psi, float ep, float mu, int bottom);
void timefield3d_set_background_permittivity(timefield3d psi, float ep);
void timefield3d_set_background_permeability(timefield3d psi, float mu);
void timefield3d_set_background_electric_conductivity(timefield3d psi, float sig);
void timefield3d_set_background_magnetic_conductivity(timefield3d psi, float sig);
void timefield3d_set_background_complex_permittivity(timefield3d psi, double omega, float real_ep, float imag_ep);
void timefield3d_set_prism_permittivity(timefield3d psi, float ep, int start_x, int end_x, int start_y, int end_y,
                                        int start_z, int end_z);
void timefield3d_set_prism_electric_conductivity(timefield3d psi, float sig, int start_x, int end_x, int start_y,
                                                 int end_y, int start_z, int end_z);
void timefield3d_set_prism_complex_permittivity(timefield3d psi, double omega, float real_ep, float imag_ep,
                                                int start_x, int end_x, int start_y, int end_y, int start_z,
                                                int end_z);
void timefield3d_set_waveguide(timefield3d psi, float ep, int start_x, int end_x, int start_z, int end_z, int bottom,
                               int top);
void timefield3d_set_taper(timefield3d psi, double ep, double mu, int start_left, int start_right, int end_left,
                           int end_right, int start_z, int end_z, int bottom, int top);

/* time stepping on the device (timefield3d.c:1537-2849).  The *_timestep pair uses explicit
 * conductivity (timefield3d.c:1537-1677); the _zsection pair and the step drivers use the
 * semi-implicit form (timefield3d.c:1679-1837).  Step drivers: snapshot of the 2-cell boundary
 * shells, E interior, E Mur faces, H interior, H Mur faces (timefield3d.c:2723-2849). */
void timefield3d_apply_H_timestep(timefield3d psi, double delta_t);
void timefield3d_apply_E_timestep(timefield3d psi, double delta_t);
void timefield3d_apply_H_timestep_zsection(timefield3d psi, double delta_t, int start_index, int stop_index);
void timefield3d_apply_E_timestep_zsection(timefield3d psi, double delta_t, int start_index, int stop_index);
void timefield3d_apply_H_first_order_abc(timefield3d psi, double *old_field, double delta_t);
void timefield3d_apply_E_first_order_abc(timefield3d psi, double *old_field, double delta_t);
void timefield3d_apply_H_first_order_abc_zsection(timefield3d psi, double delta_t, int start_index, int stop_index);
void timefield3d_apply_E_first_order_abc_zsection(timefield3d psi, double delta_t, int start_index, int stop_index);
void timefield3d_apply_timestep(timefield3d psi, double delta_t);
void timefield3d_apply_timestep_omp(timefield3d psi, double delta_t);
void timefield3d_apply_timestep_omp_full(timefield3d psi, double delta_t);

/* single-frequency running DFT into fourier_fields, F += f e^{i phase} with the reference's
 * phase reduction and series exponential (timefield3d.c:2851-2975); accumulated in FP64 on the
 * device. */
void timefield3d_update_fourier_fields(timefield3d psi, double phase);
void timefield3d_update_fourier_fields_y_slice(timefield3d psi, double phase, int slice_index);
void timefield3d_update_fourier_fields_zsection(timefield3d psi, double phase, int start_index, int stop_index);
void timefield3d_update_fourier_fields_y_slice_zsection(timefield3d psi, double phase, int slice_index,
                                                        int start_index, int stop_index);
void timefield3d_update_fourier_fields_omp(timefield3d psi, double phase);
void timefield3d_update_fourier_fields_y_slice_omp(timefield3d psi, double phase, int slice_index);

/* sources on the z = center_z plane (timefield3d.c:2977-3174) */
void timefield3d_mode_source(timefield3d psi, int center_x, int center_y, int center_z, mode source_mode,
                             double phase, double amplitude, int reverse);
void timefield3d_mode_source_adjoint(timefield3d psi, int center_x, int center_y, int center_z, mode source_mode,
                                     double phase, int reverse);
void timefield3d_TE_gaussian_source(timefield3d psi, int center_x, int center_y, int center_z, double width,
                                    double omega, double phase, double amplitude);
void timefield3d_TM_gaussian_source(timefield3d psi, int center_x, int center_y, int center_z, double width,
                                    double omega, double phase, double amplitude);

/* mode overlap on the z = center_z plane, FP64 device reduction; returns the instantaneous
 * overlap and adds e^{i phase} times it to *cumulative_overlap (timefield3d.c:3176-3239). */
complex timefield3d_mode_overlap(timefield3d psi, int center_x, int center_y, int center_z, mode overlap_mode,
                                 double phase, int reverse, complex *cumulative_overlap);

/* binary permittivity map (timefield3d.c:3241-3284) */
void timefield3d_file_write_permittivity_y_slice(timefield3d psi, char *filename, int left, int right, int start,
                                                 int end, int y_slice, double ep_min, double ep_max);

#ifdef __cplusplus
}
#endif
#endif
