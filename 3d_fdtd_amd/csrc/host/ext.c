/* Extension API (fdtd3d_ext.h): host<->device synchronisation, timers, multi-step runs, and the
 * step dispatch used by timefield3d_apply_timestep_omp_full. */
#include "fdtd3d_ext.h"

#include <string.h>

#include "fdtd3d_dev.h"
#include "host_internal.h"

struct ext_state {
    int boundary; /* reserved for boundary kinds beyond the reference's Mur faces */
};

void ext_release(tf_record *r) {
    free(r->ext);
    r->ext = NULL;
}

void tf_step(timefield3d psi, double dt) { dev_check(fdtd_dev_step(tf_dev(psi), dt, 0), "apply_timestep_omp_full"); }

const char *fdtd3d_ext_last_error(void) { return fdtd_dev_last_error(); }

int fdtd3d_ext_sync_to_device(timefield3d psi) {
    tf_lookup(psi)->mats_dirty = 1;
    fdtd_dev *d = tf_dev(psi); /* creates the device side if needed and uploads materials */
    int rc = fdtd_dev_set_fields(d, psi.fields);
    if (rc) return rc;
    return fdtd_dev_set_sources(d, psi.sources);
}

int fdtd3d_ext_sync_from_device(timefield3d psi) {
    fdtd_dev *d = tf_dev(psi);
    int rc = fdtd_dev_get_fields(d, psi.fields);
    if (rc) return rc;
    rc = fdtd_dev_get_sources(d, psi.sources);
    if (rc) return rc;
    return fdtd_dev_get_fourier(d, (double *)psi.fourier_fields);
}

int fdtd3d_ext_upload_source_plane(timefield3d psi, int k) {
    if (k < 0 || k >= psi.k) return -2;
    return fdtd_dev_set_source_plane(tf_dev(psi), k, psi.sources + 6LL * psi.i * psi.j * k, NULL);
}

int fdtd3d_ext_timer_start(timefield3d psi) { return fdtd_dev_timer_start(tf_dev(psi)); }

int fdtd3d_ext_timer_stop(timefield3d psi, double *elapsed_ms) {
    return fdtd_dev_timer_stop(tf_dev(psi), elapsed_ms);
}

int fdtd3d_ext_synchronize(timefield3d psi) { return fdtd_dev_synchronize(tf_dev(psi)); }

int fdtd3d_ext_kernel_timing(timefield3d psi, int every) { return fdtd_dev_kernel_timing(tf_dev(psi), every); }

int fdtd3d_ext_kernel_times(timefield3d psi, int counts[3], double ms[3]) {
    return fdtd_dev_kernel_times(tf_dev(psi), counts, ms);
}

int fdtd3d_ext_count_nonfinite(timefield3d psi, long long *count) {
    return fdtd_dev_count_nonfinite(tf_dev(psi), count);
}

int fdtd3d_ext_set_finite_check(timefield3d psi, int every) { return fdtd_dev_set_finite_check(tf_dev(psi), every); }

int fdtd3d_ext_device_materials(timefield3d psi, float *ep, float *mu, float *sig) {
    return fdtd_dev_get_materials(tf_dev(psi), ep, mu, sig);
}

int fdtd3d_ext_material_counts(timefield3d psi, long long counts[3]) {
    return fdtd_dev_material_counts(tf_dev(psi), counts);
}

int fdtd3d_ext_set_step_kernel(timefield3d psi, int rows, int kchunk) {
    return fdtd_dev_set_step_kernel(tf_dev(psi), rows, kchunk);
}

int fdtd3d_ext_steps(timefield3d psi, double dt, int n) {
    for (int s = 0; s < n; s++) tf_step(psi, dt);
    return 0;
}

int fdtd3d_ext_set_cpml(timefield3d psi, const int npml[6], int order, double R0, double kappa_max,
                        double alpha_max) {
    return fdtd_dev_cpml_on(tf_dev(psi), npml, order, R0, kappa_max, alpha_max);
}

int fdtd3d_ext_set_mur(timefield3d psi) { return fdtd_dev_cpml_off(tf_dev(psi)); }

int fdtd3d_ext_run_clear(timefield3d psi) { return fdtd_dev_run_clear(tf_dev(psi)); }

/* Power-normalised copy of a mode's fields (interleaved complex), as mode_power_normalize
 * would leave them, without touching the caller's mode. */
static double *normalised_copy(mode m) {
    mode c = mode_copy(m);
    mode_power_normalize(c);
    size_t n = (size_t)12 * m.i * m.j;
    double *out = malloc(n * sizeof(double));
    memcpy(out, c.fields.m, n * sizeof(double));
    mode_free(c);
    return out;
}

int fdtd3d_ext_run_add_mode_source(timefield3d psi, int cx, int cy, int cz, mode m, int reverse, int adjoint,
                                   int nsteps, const double *amplitude, const double *phase) {
    fdtd_dev *d = tf_dev(psi);
    double *f = normalised_copy(m);
    int rc = fdtd_dev_run_add_source(d, cx, cy, cz, f, m.i, m.j, m.disc, reverse, adjoint, nsteps, amplitude, phase);
    free(f);
    return rc;
}

int fdtd3d_ext_run_add_overlap(timefield3d psi, int cx, int cy, int cz, mode m, int reverse, int nfreq,
                               const double *omega) {
    fdtd_dev *d = tf_dev(psi);
    double *f = normalised_copy(m);
    int rc = fdtd_dev_run_add_overlap(d, cx, cy, cz, f, m.i, m.j, m.disc, reverse, nfreq, omega);
    free(f);
    return rc;
}

int fdtd3d_ext_run_add_dft_plane(timefield3d psi, int k, int comp_mask, int nfreq, const double *omega) {
    return fdtd_dev_run_add_dft_plane(tf_dev(psi), k, comp_mask, nfreq, omega);
}

int fdtd3d_ext_run_add_dft_box(timefield3d psi, const int box[6], int comp_mask, int nfreq, const double *omega) {
    return fdtd_dev_run_add_dft_box(tf_dev(psi), box, comp_mask, nfreq, omega);
}

int fdtd3d_ext_run_add_adjoint_source(timefield3d psi, int cx, int cy, int cz, mode m, int reverse, double omega_dt,
                                      int nsteps, const double *amplitude, const double *phase) {
    fdtd_dev *d = tf_dev(psi);
    double *f = normalised_copy(m);
    int rc = fdtd_dev_run_add_adjoint_source(d, cx, cy, cz, f, m.i, m.j, m.disc, reverse, omega_dt, nsteps, amplitude,
                                             phase);
    free(f);
    return rc;
}

int fdtd3d_ext_dft_gradient(timefield3d psi_f, int id_f, timefield3d psi_a, int id_a, int freq, complex w,
                            double *grad) {
    return fdtd_dev_dft_gradient(tf_dev(psi_f), id_f, tf_dev(psi_a), id_a, freq, w.re, w.im, grad);
}

int fdtd3d_ext_run(timefield3d psi, double dt, int n0, int count, double t_offset) {
    return fdtd_dev_run(tf_dev(psi), dt, n0, count, t_offset);
}

int fdtd3d_ext_run_overlap(timefield3d psi, int id, complex *cumulative, complex *last) {
    return fdtd_dev_run_overlap_result(tf_dev(psi), id, (double *)cumulative, (double *)last);
}

int fdtd3d_ext_run_dft(timefield3d psi, int id, complex *out) {
    return fdtd_dev_run_dft_result(tf_dev(psi), id, (double *)out);
}
