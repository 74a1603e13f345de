/*
 * fdtd3d_ext.h -- extensions beyond the reference C surface (no counterpart in the reference).
 *
 * For callers that need more than the drop-in API: explicit host<->device synchronisation of
 * the timefield3d mirrors, timers on the grid's HIP stream, and the error text of the last
 * failing device call.  Every function returns 0 on success and a negative code on failure.
 */
#ifndef FDTD3D_EXT_H
#define FDTD3D_EXT_H

#include "timefield3d.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Upload ep/mu/sig, fields and sources from the host arrays of psi (after direct writes). */
int fdtd3d_ext_sync_to_device(timefield3d psi);
/* Refresh the host arrays fields, sources and fourier_fields of psi from the device. */
int fdtd3d_ext_sync_from_device(timefield3d psi);
/* Upload z plane k of the host sources array only (after direct writes to psi.sources). */
int fdtd3d_ext_upload_source_plane(timefield3d psi, int k);

/* Time the work enqueued on psi's stream between start and stop (HIP events on that stream). */
int fdtd3d_ext_timer_start(timefield3d psi);
int fdtd3d_ext_timer_stop(timefield3d psi, double *elapsed_ms);
int fdtd3d_ext_synchronize(timefield3d psi);
/* Sample every `every`-th launch of each Yee update kernel with HIP events (0 switches
 * sampling off); kernel_times returns per slot (0 split E, 1 split H, 2 fused step) the number
 * of sampled launches and their summed durations in ms. */
int fdtd3d_ext_kernel_timing(timefield3d psi, int every);
int fdtd3d_ext_kernel_times(timefield3d psi, int counts[3], double ms[3]);
/* Non-finite field detection (no counterpart in the reference): count the NaN / Inf field values
 * now (synchronous), or let fdtd3d_ext_run scan every `every` steps and fail with -4 (0 = off);
 * see fdtd_dev_count_nonfinite / fdtd_dev_set_finite_check. */
int fdtd3d_ext_count_nonfinite(timefield3d psi, long long *count);
int fdtd3d_ext_set_finite_check(timefield3d psi, int every);

/* The device's copy of the materials (no counterpart in the reference): per-cell AoS ep, mu, sig
 * (NULL skips one) expanded from the device ids, and counts[0] full material uploads, counts[1]
 * setter calls applied on the device (timefield3d_set_taper, _set_prism_permittivity,
 * _set_waveguide and adjoint3d_taper_insert repaint the device ids when the device materials are
 * current), counts[2] material table entries. */
int fdtd3d_ext_device_materials(timefield3d psi, float *ep, float *mu, float *sig);
int fdtd3d_ext_material_counts(timefield3d psi, long long counts[3]);
/* Under the PML boundary: rows = 0 selects the split E/H kernels (kchunk > 0: z planes per
 * thread, 0: automatic), rows = -1 the fused one-kernel step in columns (kchunk = 1, 2 or 4 z
 * planes per thread), rows = 1, 2, 4 or 8 the fused step in register strips (rows per lane,
 * kchunk z planes per wavefront), rows = 16 or 32 the fused step in LDS tiles of 8 or 16 rows
 * (kchunk z planes per block).  Default: split kernels, automatic chunking. */
int fdtd3d_ext_set_step_kernel(timefield3d psi, int rows, int kchunk);

/* Run n reference steps (timefield3d_apply_timestep_omp_full semantics) without host work in
 * between; sources stay as last set. */
int fdtd3d_ext_steps(timefield3d psi, double delta_t, int n);

/* Boundary.  The reference terminates the grid with first-order Mur faces (the default).
 * fdtd3d_ext_set_cpml switches every later step of psi -- including
 * timefield3d_apply_timestep_omp_full -- to a convolutional PML (Roden-Gedney, CFS grading)
 * of npml[6] cells on faces x-lo, x-hi, y-lo, y-hi, z-lo, z-hi in front of PEC walls:
 * sigma = -(order+1) ln(R0) / (2 npml dx) rho^order, kappa = 1 + (kappa_max-1) rho^order,
 * alpha = alpha_max (1 - rho).  fdtd3d_ext_set_mur restores the reference boundary. */
int fdtd3d_ext_set_cpml(timefield3d psi, const int npml[6], int order, double R0, double kappa_max,
                        double alpha_max);
int fdtd3d_ext_set_mur(timefield3d psi);

/* Device-resident runs: sources with tabulated amplitude/phase, FP64 overlap monitors with a
 * running DFT at several angular frequencies, FP64 multi-frequency DFT planes.  Step n of
 * fdtd3d_ext_run(psi, dt, n0, count, t_offset) writes each source as the reference's
 * timefield3d_mode_source(_adjoint) would with amplitude[n] and phase[n] (a power-normalised copy
 * of the mode is used; the caller's mode is not modified), advances one step, then adds
 * e^{i omega (n + t_offset) dt} times the instantaneous overlap (or the field) to each monitor.
 * Sources are off outside [0, nsteps).  The add functions return a non-negative id. */
int fdtd3d_ext_run_clear(timefield3d psi);
int fdtd3d_ext_run_add_mode_source(timefield3d psi, int center_x, int center_y, int center_z, mode source_mode,
                                   int reverse, int adjoint, int nsteps, const double *amplitude,
                                   const double *phase);
int fdtd3d_ext_run_add_overlap(timefield3d psi, int center_x, int center_y, int center_z, mode overlap_mode,
                               int reverse, int nfreq, const double *omega);
/* comp_mask bit m selects field component m (Ex=bit0 ... Hz=bit5) on plane z = k. */
int fdtd3d_ext_run_add_dft_plane(timefield3d psi, int k, int comp_mask, int nfreq, const double *omega);
int fdtd3d_ext_run(timefield3d psi, double delta_t, int n0, int count, double t_offset);
/* cumulative[nfreq] and the last instantaneous overlap (either pointer may be NULL). */
int fdtd3d_ext_run_overlap(timefield3d psi, int id, complex *cumulative, complex *last);
/* DFT monitor over the box [box[0],box[1]) x [box[2],box[3]) x [box[4],box[5]) in (i, j, k); shares
 * the id space of the DFT planes. */
int fdtd3d_ext_run_add_dft_box(timefield3d psi, const int box[6], int comp_mask, int nfreq, const double *omega);
/* [selected comp][freq][j][i] complex values of a DFT plane ([comp][freq][k][j][i] of a box). */
int fdtd3d_ext_run_dft(timefield3d psi, int id, complex *out);

/* DFT-domain permittivity gradient.
 * Adjoint source of an overlap monitor a(omega) = sum_n e^{i omega (n + t_offset) dt} inst(n):
 * fdtd3d_ext_run_add_adjoint_source places, on the monitor plane, the electric and magnetic
 * currents of the monitor's weights, J = dx^2/2 (mHy', -mHx', 0) and
 * M = e^{-i omega_dt} dx^2/2 (mEy, -mEx, 0) (mH' = -conj(mH) when reverse), written at step n as
 * Re[(J, M) amplitude[n] e^{-i phase[n]}] -- the reference's CW adjoint source form
 * (timefield3d.c:3028-3068) with the scale and phase that make the discrete adjoint exact.
 * Forward run: the monitor and a DFT box of E with t_offset 0.  Adjoint run (fresh fields, same
 * materials): this source with phase[n] = omega n dt and a smooth envelope amplitude[n], and a DFT
 * box of E with t_offset 1.  Then, with u = e^{-i omega dt} - 1 and E0 = sum_n amplitude[n],
 *   d a / d eps(cell) = 2 u / (dt E0) * sum_m F_m(cell) A_m(cell)
 * for a change of eps on all three components of the cell (first order, fields decayed at the end
 * of both runs, sources and monitor outside the box).  fdtd3d_ext_dft_gradient returns
 * grad[cell] = Re[w * sum_m F_m A_m] over the box (x fastest) for a complex weight w, e.g.
 * w = 2 conj(a) 2u/(dt E0) for d|a|^2/d eps. */
int fdtd3d_ext_run_add_adjoint_source(timefield3d psi, int center_x, int center_y, int center_z, mode overlap_mode,
                                      int reverse, double omega_dt, int nsteps, const double *amplitude,
                                      const double *phase);
int fdtd3d_ext_dft_gradient(timefield3d psi_forward, int id_forward, timefield3d psi_adjoint, int id_adjoint,
                            int freq, complex weight, double *grad);

const char *fdtd3d_ext_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
