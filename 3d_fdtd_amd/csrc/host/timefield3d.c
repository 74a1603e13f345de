/* Host side of the timefield3d drop-in surface (timefield3d.h).
 *
 * Every grid owns device state (fdtd3d_dev.h) found through a registry keyed by the host
 * fields pointer, so by-value copies of the struct share it.  ep/mu/sig are host-authoritative
 * (setters mark the device copy stale); fields, sources and fourier_fields are
 * device-authoritative and refreshed into the host arrays by the calls that read them.
 * Reference behaviour cited per function (timefield3d.c line numbers). */
#include "timefield3d.h"

#include <string.h>

#include "fdtd3d_dev.h"
#include "fdtd3d_ext.h"
#include "host_internal.h"

#define TF_PI 3.141592652 /* the reference's phase-reduction pi (timefield3d.c:3) */

/* ------------------------------------------------------------------------------------------ */
/* registry                                                                                   */
/* ------------------------------------------------------------------------------------------ */

static tf_record *g_recs = NULL;
static int g_nrec = 0, g_cap = 0;

tf_record *tf_lookup(timefield3d psi) {
    for (int q = 0; q < g_nrec; q++)
        if (g_recs[q].key == psi.fields) return &g_recs[q];
    fprintf(stderr, "fdtd3d: grid %p was not created by timefield3d_new\n", (void *)psi.fields);
    abort();
}

static tf_record *tf_register(timefield3d psi, fdtd_dev *dev) {
    if (g_nrec == g_cap) {
        g_cap = g_cap ? 2 * g_cap : 8;
        g_recs = realloc(g_recs, sizeof(tf_record) * g_cap);
    }
    tf_record *r = &g_recs[g_nrec++];
    memset(r, 0, sizeof *r);
    r->key = psi.fields;
    r->dev = dev;
    r->mats_dirty = 1;
    return r;
}

static void tf_unregister(timefield3d psi) {
    for (int q = 0; q < g_nrec; q++)
        if (g_recs[q].key == psi.fields) {
            g_recs[q] = g_recs[--g_nrec];
            return;
        }
}

void dev_check(int rc, const char *where) {
    if (rc != 0) {
        fprintf(stderr, "fdtd3d: %s failed: %s\n", where, fdtd_dev_last_error());
        abort();
    }
}

/* Device state of psi with materials current.  The device side is created on first use, so
 * host-only work (setters, file output) never touches the GPU. */
fdtd_dev *tf_dev(timefield3d psi) {
    tf_record *r = tf_lookup(psi);
    if (!r->dev) {
        r->dev = fdtd_dev_create(psi.i, psi.j, psi.k, psi.disc);
        if (!r->dev) {
            fprintf(stderr, "fdtd3d: device state for a %dx%dx%d grid: %s\n", psi.i, psi.j, psi.k,
                    fdtd_dev_last_error());
            abort();
        }
    }
    if (r->mats_dirty) {
        dev_check(fdtd_dev_set_materials(r->dev, psi.ep, psi.mu, psi.sig), "material upload");
        r->mats_dirty = 0;
    }
    return r->dev;
}

static void mark_materials(timefield3d psi) { tf_lookup(psi)->mats_dirty = 1; }

/* The device state of psi when its materials are current there (a setter can then repaint the
 * device ids itself instead of leaving the whole volume to the next upload), else NULL. */
static fdtd_dev *current_device(timefield3d psi) {
    tf_record *r = tf_lookup(psi);
    return r->dev && !r->mats_dirty ? r->dev : NULL;
}

static long long ncell(timefield3d psi) { return (long long)psi.i * psi.j * psi.k; }

static void pull_fields(timefield3d psi) { dev_check(fdtd_dev_get_fields(tf_dev(psi), psi.fields), "field download"); }

/* ------------------------------------------------------------------------------------------ */
/* lifecycle (timefield3d.c:6-346)                                                            */
/* ------------------------------------------------------------------------------------------ */

timefield3d timefield3d_new(int i, int j, int k, double disc) {
    timefield3d t;
    long long n = (long long)i * j * k;
    t.i = i;
    t.j = j;
    t.k = k;
    t.disc = disc;
    t.ep = malloc(sizeof(float) * 3 * n);
    t.mu = malloc(sizeof(float) * 3 * n);
    t.sig = calloc(6 * n, sizeof(float));
    t.fields = calloc(6 * n, sizeof(double));
    t.old_fields = calloc(6 * n, sizeof(double));
    t.fourier_fields = calloc(6 * n, sizeof(complex));
    t.sources = calloc(6 * n, sizeof(float));
    if (!t.ep || !t.mu || !t.sig || !t.fields || !t.old_fields || !t.fourier_fields || !t.sources) {
        fprintf(stderr, "fdtd3d: timefield3d_new: out of host memory\n");
        abort();
    }
    for (long long q = 0; q < 3 * n; q++) {
        t.ep[q] = 1;
        t.mu[q] = 1;
    }
    tf_register(t, NULL);
    return t;
}

void timefield3d_free(timefield3d psi) {
    tf_record *r = tf_lookup(psi);
    ext_release(r);
    if (r->dev) fdtd_dev_destroy(r->dev);
    tf_unregister(psi);
    free(psi.ep);
    free(psi.mu);
    free(psi.sig);
    free(psi.fields);
    free(psi.old_fields);
    free(psi.fourier_fields);
    free(psi.sources);
}

/* Deep copy.  Like the reference, the copy's sig[6c] receives fields[6c] instead of sig[6c]
 * (timefield3d.c:73). */
timefield3d timefield3d_copy(timefield3d psi) {
    timefield3d c = timefield3d_new(psi.i, psi.j, psi.k, psi.disc);
    long long n = ncell(psi);
    dev_check(fdtd3d_ext_sync_from_device(psi), "copy: download");
    memcpy(c.ep, psi.ep, sizeof(float) * 3 * n);
    memcpy(c.mu, psi.mu, sizeof(float) * 3 * n);
    memcpy(c.sig, psi.sig, sizeof(float) * 6 * n);
    for (long long q = 0; q < n; q++) c.sig[6 * q] = (float)psi.fields[6 * q];
    memcpy(c.fields, psi.fields, sizeof(double) * 6 * n);
    memcpy(c.old_fields, psi.old_fields, sizeof(double) * 6 * n);
    memcpy(c.fourier_fields, psi.fourier_fields, sizeof(complex) * 6 * n);
    memcpy(c.sources, psi.sources, sizeof(float) * 6 * n);
    dev_check(fdtd3d_ext_sync_to_device(c), "copy: upload");
    dev_check(fdtd_dev_set_fourier(tf_dev(c), (const double *)c.fourier_fields), "copy: fourier upload");
    return c;
}

/* Zero fields, boundary history, DFT accumulators and sources (timefield3d.c:108-121).  The
 * host mirrors are refreshed from the device when next read. */
void timefield3d_zero_fields(timefield3d psi) { dev_check(fdtd_dev_zero(tf_dev(psi)), "zero_fields"); }

double *timefield3d_get_fields(timefield3d psi) {
    long long n = 6 * ncell(psi);
    double *out = malloc(sizeof(double) * n);
    pull_fields(psi);
    memcpy(out, psi.fields, sizeof(double) * n);
    return out;
}

/* Full-size array holding only the two outermost layers of every face (timefield3d.c:138-244);
 * cells elsewhere are zero. */
double *timefield3d_get_edge_fields(timefield3d psi) {
    long long n = ncell(psi);
    double *out = calloc(6 * n, sizeof(double));
    pull_fields(psi);
    const int N[3] = {psi.i, psi.j, psi.k};
    for (int a = 0; a < 3; a++) {
        int b = (a + 1) % 3, c = (a + 2) % 3;
        for (int side = 0; side < 2; side++)
            for (int layer = 0; layer < 2; layer++)
                for (int v = 0; v < N[c]; v++)
                    for (int u = 0; u < N[b]; u++) {
                        int x[3];
                        x[a] = side == 0 ? layer : N[a] - 1 - layer;
                        x[b] = u;
                        x[c] = v;
                        long long q = x[0] + (long long)N[0] * (x[1] + (long long)N[1] * x[2]);
                        memcpy(out + 6 * q, psi.fields + 6 * q, 6 * sizeof(double));
                    }
    }
    return out;
}

void timefield3d_update_old_edge_fields(timefield3d psi) {
    dev_check(fdtd_dev_snapshot_shells(tf_dev(psi)), "update_old_edge_fields");
}

/* ------------------------------------------------------------------------------------------ */
/* plots: text slices into $FDTD3D_PLOT_DIR                                                   */
/* ------------------------------------------------------------------------------------------ */

static FILE *plot_open(const char *name, int slice) {
    const char *dir = getenv("FDTD3D_PLOT_DIR");
    if (!dir || !*dir) return NULL;
    char path[4096];
    snprintf(path, sizeof path, "%s/%s_%d.txt", dir, name, slice);
    return fopen(path, "w");
}

/* One slice as the gnuplot "nonuniform matrix" data file the reference's plot functions write
 * before piping it to gnuplot (timefield3d.c:348-1354): a header row with the column count and
 * the column coordinates, then one row per row coordinate followed by its values, every number
 * as "%lf\t".  Coordinates are cell centres relative to the grid centre, (n + 0.5 - N/2) disc.
 * x slice (plane i = s): rows j, columns k; y slice (plane j = s): rows i, columns k; z slice
 * (plane k = s): rows j, columns i.  what: 0..5 field component, 6 = |fourier Ex|, 7 = ep_x.
 * The header's first number is the true column count (the reference writes psi.k for every
 * slice, which misreads z slices whenever psi.i != psi.k). */
static void plot_slice(timefield3d psi, int axis, int s, int what, const char *name) {
    const int n_axis = axis == 0 ? psi.i : (axis == 1 ? psi.j : psi.k);
    if (s < 0 || s >= n_axis) {
        fprintf(stderr, "fdtd3d: %s slice %d outside 0..%d\n", name, s, n_axis - 1);
        return;
    }
    FILE *f = plot_open(name, s);
    if (!f) return;
    if (what < 6) pull_fields(psi);
    if (what == 6) dev_check(fdtd_dev_get_fourier(tf_dev(psi), (double *)psi.fourier_fields), "fourier download");
    const int N[3] = {psi.i, psi.j, psi.k};
    const int ra = axis == 1 ? 0 : 1;  // row axis
    const int ca = axis == 2 ? 0 : 2;  // column axis
    fprintf(f, "%lf\t", (double)N[ca]);
    for (int c = 0; c < N[ca]; c++) fprintf(f, "%lf\t", (c + 0.5 - N[ca] / 2.0) * psi.disc);
    fprintf(f, "\n");
    for (int r = 0; r < N[ra]; r++) {
        fprintf(f, "%lf\t", (r + 0.5 - N[ra] / 2.0) * psi.disc);
        for (int c = 0; c < N[ca]; c++) {
            int x[3];
            x[axis] = s;
            x[ra] = r;
            x[ca] = c;
            long long q = x[0] + (long long)psi.i * (x[1] + (long long)psi.j * x[2]);
            double v;
            if (what < 6) v = psi.fields[6 * q + what];
            else if (what == 6) v = complex_mag(psi.fourier_fields[6 * q]);
            else v = psi.ep[3 * q];
            fprintf(f, "%lf\t", v);
        }
        fprintf(f, "\n");
    }
    fclose(f);
}

void timefield3d_plot_Ex_x_slice(timefield3d psi, int s) { plot_slice(psi, 0, s, 0, "Ex_x"); }
void timefield3d_plot_Ex_y_slice(timefield3d psi, int s) { plot_slice(psi, 1, s, 0, "Ex_y"); }
void timefield3d_plot_Ex_z_slice(timefield3d psi, int s) { plot_slice(psi, 2, s, 0, "Ex_z"); }
void timefield3d_plot_Ey_x_slice(timefield3d psi, int s) { plot_slice(psi, 0, s, 1, "Ey_x"); }
void timefield3d_plot_Ey_y_slice(timefield3d psi, int s) { plot_slice(psi, 1, s, 1, "Ey_y"); }
void timefield3d_plot_Ey_z_slice(timefield3d psi, int s) { plot_slice(psi, 2, s, 1, "Ey_z"); }
void timefield3d_plot_Ez_x_slice(timefield3d psi, int s) { plot_slice(psi, 0, s, 2, "Ez_x"); }
void timefield3d_plot_Ez_y_slice(timefield3d psi, int s) { plot_slice(psi, 1, s, 2, "Ez_y"); }
void timefield3d_plot_Ez_z_slice(timefield3d psi, int s) { plot_slice(psi, 2, s, 2, "Ez_z"); }
void timefield3d_plot_Ex_x_slice_fourier(timefield3d psi, int s) { plot_slice(psi, 0, s, 6, "Ex_fourier_x"); }
void timefield3d_plot_Ex_y_slice_fourier(timefield3d psi, int s) { plot_slice(psi, 1, s, 6, "Ex_fourier_y"); }
void timefield3d_plot_Ex_z_slice_fourier(timefield3d psi, int s) { plot_slice(psi, 2, s, 6, "Ex_fourier_z"); }
void timefield3d_plot_permittivity_x_slice(timefield3d psi, int s) { plot_slice(psi, 0, s, 7, "ep_x"); }
void timefield3d_plot_permittivity_y_slice(timefield3d psi, int s) { plot_slice(psi, 1, s, 7, "ep_y"); }
void timefield3d_plot_permittivity_z_slice(timefield3d psi, int s) { plot_slice(psi, 2, s, 7, "ep_z"); }

/* ------------------------------------------------------------------------------------------ */
/* material setters (timefield3d.c:1356-1535)                                                 */
/* ------------------------------------------------------------------------------------------ */

void timefield3d_set_background(timefield3d psi, float ep, float mu) {
    long long n = 3 * ncell(psi);
    for (long long q = 0; q < n; q++) {
        psi.ep[q] = ep;
        psi.mu[q] = mu;
    }
    mark_materials(psi);
}

void timefield3d_set_top_half_plane(timefield3d psi, float ep, float mu, int bottom) {
    for (int k = 0; k < psi.k; k++)
        for (int j = bottom < 0 ? 0 : bottom; j < psi.j; j++)
            for (int i = 0; i < psi.i; i++) {
                long long q = 3 * (i + (long long)psi.i * (j + (long long)psi.j * k));
                for (int m = 0; m < 3; m++) {
                    psi.ep[q + m] = ep;
                    psi.mu[q + m] = mu;
                }
            }
    mark_materials(psi);
}

void timefield3d_set_background_permittivity(timefield3d psi, float ep) {
    for (long long q = 0; q < 3 * ncell(psi); q++) psi.ep[q] = ep;
    mark_materials(psi);
}

void timefield3d_set_background_permeability(timefield3d psi, float mu) {
    for (long long q = 0; q < 3 * ncell(psi); q++) psi.mu[q] = mu;
    mark_materials(psi);
}

void timefield3d_set_background_electric_conductivity(timefield3d psi, float sig) {
    for (long long q = 0; q < ncell(psi); q++)
        for (int m = 0; m < 3; m++) psi.sig[6 * q + m] = sig;
    mark_materials(psi);
}

void timefield3d_set_background_magnetic_conductivity(timefield3d psi, float sig) {
    for (long long q = 0; q < ncell(psi); q++)
        for (int m = 3; m < 6; m++) psi.sig[6 * q + m] = sig;
    mark_materials(psi);
}

/* Im(ep) enters as an electric conductivity omega*Im(ep), rounded to float (timefield3d.c:1445). */
void timefield3d_set_background_complex_permittivity(timefield3d psi, double omega, float real_ep, float imag_ep) {
    timefield3d_set_background_permittivity(psi, real_ep);
    timefield3d_set_background_electric_conductivity(psi, omega * imag_ep);
}

static void clip_range(int *lo, int *hi, int n) {
    if (*lo < 0) *lo = 0;
    if (*hi > n) *hi = n;
}

void timefield3d_set_prism_permittivity(timefield3d psi, float ep, int x0, int x1, int y0, int y1, int z0, int z1) {
    clip_range(&x0, &x1, psi.i);
    clip_range(&y0, &y1, psi.j);
    clip_range(&z0, &z1, psi.k);
    for (int k = z0; k < z1; k++)
        for (int j = y0; j < y1; j++)
            for (int i = x0; i < x1; i++) {
                long long q = 3 * (i + (long long)psi.i * (j + (long long)psi.j * k));
                psi.ep[q] = psi.ep[q + 1] = psi.ep[q + 2] = ep;
            }
    fdtd_dev *d = current_device(psi);
    if (d)
        dev_check(fdtd_dev_paint_prism(d, ep, 1.0f, 0, x0, x1, y0, y1, z0, z1), "prism on the device");
    else
        mark_materials(psi);
}

void timefield3d_set_prism_electric_conductivity(timefield3d psi, float sig, int x0, int x1, int y0, int y1, int z0,
                                                 int z1) {
    clip_range(&x0, &x1, psi.i);
    clip_range(&y0, &y1, psi.j);
    clip_range(&z0, &z1, psi.k);
    for (int k = z0; k < z1; k++)
        for (int j = y0; j < y1; j++)
            for (int i = x0; i < x1; i++) {
                long long q = 6 * (i + (long long)psi.i * (j + (long long)psi.j * k));
                psi.sig[q] = psi.sig[q + 1] = psi.sig[q + 2] = sig;
            }
    mark_materials(psi);
}

void timefield3d_set_prism_complex_permittivity(timefield3d psi, double omega, float real_ep, float imag_ep, int x0,
                                                int x1, int y0, int y1, int z0, int z1) {
    timefield3d_set_prism_permittivity(psi, real_ep, x0, x1, y0, y1, z0, z1);
    timefield3d_set_prism_electric_conductivity(psi, omega * imag_ep, x0, x1, y0, y1, z0, z1);
}

void timefield3d_set_waveguide(timefield3d psi, float ep, int start_x, int end_x, int start_z, int end_z, int bottom,
                               int top) {
    timefield3d_set_prism_permittivity(psi, ep, start_x, end_x, bottom, top, start_z, end_z);
}

/* Linear taper: row k spans [l0 + round(sl*(k-z0)), r0 + round(sr*(k-z0))) in x for y in
 * [bottom, top); ep and mu of all three components, stored as float (timefield3d.c:1511-1535). */
void timefield3d_set_taper(timefield3d psi, double ep, double mu, int start_left, int start_right, int end_left,
                           int end_right, int start_z, int end_z, int bottom, int top) {
    double rs = (double)(end_right - start_right) / ((double)(end_z - start_z));
    double ls = (double)(end_left - start_left) / ((double)(end_z - start_z));
    for (int j = bottom; j < top; j++)
        for (int k = start_z; k < end_z; k++) {
            int lo = (int)(start_left + round(ls * (k - start_z)));
            double hi = start_right + round(rs * (k - start_z));
            for (int i = lo; i < hi; i++) {
                long long q = 3 * (i + (long long)psi.i * (j + (long long)psi.j * k));
                for (int m = 0; m < 3; m++) {
                    psi.ep[q + m] = (float)ep;
                    psi.mu[q + m] = (float)mu;
                }
            }
        }
    fdtd_dev *d = current_device(psi);
    if (d)
        dev_check(fdtd_dev_paint_taper(d, (float)ep, (float)mu, start_left, start_right, end_left, end_right, start_z,
                                       end_z, bottom, top),
                  "taper on the device");
    else
        mark_materials(psi);
}

/* ------------------------------------------------------------------------------------------ */
/* time stepping (timefield3d.c:1537-2849)                                                    */
/* ------------------------------------------------------------------------------------------ */

enum { SEMI = 0, EXPLICIT = 1 };

void timefield3d_apply_H_timestep(timefield3d psi, double dt) {
    dev_check(fdtd_dev_update_H(tf_dev(psi), dt, 1, psi.k - 1, EXPLICIT), "apply_H_timestep");
}

void timefield3d_apply_E_timestep(timefield3d psi, double dt) {
    dev_check(fdtd_dev_update_E(tf_dev(psi), dt, 1, psi.k - 1, EXPLICIT), "apply_E_timestep");
}

/* z range [start+1, stop-1) as in the reference slab loops (timefield3d.c:1693,1773). */
void timefield3d_apply_H_timestep_zsection(timefield3d psi, double dt, int start, int stop) {
    dev_check(fdtd_dev_update_H(tf_dev(psi), dt, start + 1, stop - 1, SEMI), "apply_H_timestep_zsection");
}

void timefield3d_apply_E_timestep_zsection(timefield3d psi, double dt, int start, int stop) {
    dev_check(fdtd_dev_update_E(tf_dev(psi), dt, start + 1, stop - 1, SEMI), "apply_E_timestep_zsection");
}

void timefield3d_apply_H_first_order_abc(timefield3d psi, double *old_field, double dt) {
    fdtd_dev *d = tf_dev(psi);
    dev_check(fdtd_dev_set_shells(d, old_field), "abc: shell upload");
    dev_check(fdtd_dev_mur_H(d, dt, 0, psi.k), "apply_H_first_order_abc");
}

void timefield3d_apply_E_first_order_abc(timefield3d psi, double *old_field, double dt) {
    fdtd_dev *d = tf_dev(psi);
    dev_check(fdtd_dev_set_shells(d, old_field), "abc: shell upload");
    dev_check(fdtd_dev_mur_E(d, dt, 0, psi.k), "apply_E_first_order_abc");
}

void timefield3d_apply_H_first_order_abc_zsection(timefield3d psi, double dt, int start, int stop) {
    dev_check(fdtd_dev_mur_H(tf_dev(psi), dt, start, stop), "apply_H_first_order_abc_zsection");
}

void timefield3d_apply_E_first_order_abc_zsection(timefield3d psi, double dt, int start, int stop) {
    dev_check(fdtd_dev_mur_E(tf_dev(psi), dt, start, stop), "apply_E_first_order_abc_zsection");
}

/* Serial driver: explicit conductivity (timefield3d.c:2723-2735). */
void timefield3d_apply_timestep(timefield3d psi, double dt) {
    dev_check(fdtd_dev_step(tf_dev(psi), dt, EXPLICIT), "apply_timestep");
}

/* Both OpenMP drivers use the semi-implicit slab updates (timefield3d.c:2737-2849). */
void timefield3d_apply_timestep_omp(timefield3d psi, double dt) {
    dev_check(fdtd_dev_step(tf_dev(psi), dt, SEMI), "apply_timestep_omp");
}

void timefield3d_apply_timestep_omp_full(timefield3d psi, double dt) { tf_step(psi, dt); }

/* ------------------------------------------------------------------------------------------ */
/* running DFT (timefield3d.c:2851-2975)                                                      */
/* ------------------------------------------------------------------------------------------ */

/* e^{i phase~}, phase~ = phase - (int(phase/2pi~)*2)*pi~, series exponential. */
void tf_twiddle(double phase, double *re, double *im) {
    int turns = (int)(phase / (2 * TF_PI));
    double reduced = phase - (turns * 2) * TF_PI;
    complex t = complex_exp(complex_from_doubles(0, reduced));
    *re = t.re;
    *im = t.im;
}

static void dft(timefield3d psi, double phase, int y, int k0, int k1) {
    double re, im;
    tf_twiddle(phase, &re, &im);
    dev_check(fdtd_dev_dft(tf_dev(psi), re, im, y, k0, k1), "update_fourier_fields");
}

void timefield3d_update_fourier_fields(timefield3d psi, double phase) { dft(psi, phase, -1, 0, psi.k); }
void timefield3d_update_fourier_fields_y_slice(timefield3d psi, double phase, int s) { dft(psi, phase, s, 0, psi.k); }
void timefield3d_update_fourier_fields_zsection(timefield3d psi, double phase, int start, int stop) {
    dft(psi, phase, -1, start, stop);
}
void timefield3d_update_fourier_fields_y_slice_zsection(timefield3d psi, double phase, int s, int start, int stop) {
    dft(psi, phase, s, start, stop);
}
void timefield3d_update_fourier_fields_omp(timefield3d psi, double phase) { dft(psi, phase, -1, 0, psi.k); }
void timefield3d_update_fourier_fields_y_slice_omp(timefield3d psi, double phase, int s) {
    dft(psi, phase, s, 0, psi.k);
}

/* ------------------------------------------------------------------------------------------ */
/* sources (timefield3d.c:2977-3174) and overlap (timefield3d.c:3176-3239)                    */
/* ------------------------------------------------------------------------------------------ */

void timefield3d_mode_source(timefield3d psi, int cx, int cy, int cz, mode m, double phase, double amplitude,
                             int reverse) {
    mode_power_normalize(m);
    dev_check(fdtd_dev_mode_source(tf_dev(psi), cx, cy, cz, (const double *)m.fields.m, m.i, m.j, m.disc, cos(phase),
                                   sin(phase), amplitude, reverse, 0),
              "mode_source");
}

void timefield3d_mode_source_adjoint(timefield3d psi, int cx, int cy, int cz, mode m, double phase, int reverse) {
    mode_power_normalize(m);
    dev_check(fdtd_dev_mode_source(tf_dev(psi), cx, cy, cz, (const double *)m.fields.m, m.i, m.j, m.disc, cos(phase),
                                   sin(phase), 1.0, reverse, 1),
              "mode_source_adjoint");
}

/* Gaussian beam source: the mode comes from mode_gaussian with the complex ep/mu of the centre
 * cell, and each plane cell uses its own complex ep/mu (timefield3d.c:3070-3174).  The window
 * test is on the flattened mode index, as in the reference. */
static void gaussian_source(timefield3d psi, int cx, int cy, int cz, double width, double omega, double phase,
                            double amplitude, int polarization) {
    long long c0 = cx + (long long)psi.i * (cy + (long long)psi.j * cz);
    complex ep = complex_from_doubles(psi.ep[3 * c0], psi.sig[6 * c0] / omega);
    complex mu = complex_from_doubles(psi.mu[3 * c0], psi.sig[6 * c0 + 3] / omega);
    int size = (int)(3.0 * width / psi.disc);
    mode g = mode_gaussian(size, size, psi.disc, ep, mu, omega, (width / (2.35482 * psi.disc)), polarization);
    long long P = (long long)psi.i * psi.j;
    float *plane = calloc(6 * P, sizeof(float));
    unsigned char *mask = calloc(P, 1);
    for (int i = 0; i < psi.i; i++)
        for (int j = 0; j < psi.j; j++) {
            long long q = i + (long long)psi.i * (j + (long long)psi.j * cz);
            long long p = i + (long long)psi.i * j;
            long long mi = 6LL * ((i - cx + size / 2) + (long long)g.i * (j - cy + size / 2));
            if (mi < 0 || mi >= 6LL * g.i * g.j) continue;
            mask[p] = 1;
            for (int m = 0; m < 3; m++) {
                complex e = complex_from_doubles(psi.ep[3 * q + m], psi.sig[6 * q + m] / omega);
                complex u = complex_from_doubles(psi.mu[3 * q + m], psi.sig[6 * q + m + 3] / omega);
                complex fe = g.fields.m[mi + m], fh = g.fields.m[mi + m + 3];
                double sre = amplitude * (e.re * fe.re - e.im * fe.im);
                double sim = amplitude * (e.re * fe.im + e.im * fe.re);
                plane[6 * p + m] = (float)(sre * cos(phase) + sim * sin(phase));
                sre = amplitude * (u.re * fh.re - u.im * fh.im);
                sim = amplitude * (u.re * fh.im + u.im * fh.re);
                plane[6 * p + m + 3] = (float)(sre * cos(phase) + sim * sin(phase));
            }
        }
    dev_check(fdtd_dev_set_source_plane(tf_dev(psi), cz, plane, mask), "gaussian source");
    free(plane);
    free(mask);
    mode_free(g);
}

void timefield3d_TE_gaussian_source(timefield3d psi, int cx, int cy, int cz, double width, double omega, double phase,
                                    double amplitude) {
    gaussian_source(psi, cx, cy, cz, width, omega, phase, amplitude, 0);
}

void timefield3d_TM_gaussian_source(timefield3d psi, int cx, int cy, int cz, double width, double omega, double phase,
                                    double amplitude) {
    gaussian_source(psi, cx, cy, cz, width, omega, phase, amplitude, 1);
}

complex timefield3d_mode_overlap(timefield3d psi, int cx, int cy, int cz, mode m, double phase, int reverse,
                                 complex *cumulative) {
    mode_power_normalize(m);
    double out[2];
    dev_check(fdtd_dev_mode_overlap(tf_dev(psi), cx, cy, cz, (const double *)m.fields.m, m.i, m.j, m.disc, reverse,
                                    out),
              "mode_overlap");
    complex inst = complex_from_doubles(out[0], out[1]);
    double tr, ti;
    tf_twiddle(phase, &tr, &ti);
    *cumulative = complex_add(*cumulative, complex_mul(complex_from_doubles(tr, ti), inst));
    return inst;
}

/* Header int y, int z, double disc, then int[y*z] with 1 where ep_x > (ep_min+ep_max)/2
 * (timefield3d.c:3241-3284). */
void timefield3d_file_write_permittivity_y_slice(timefield3d psi, char *filename, int left, int right, int start,
                                                 int end, int y_slice, double ep_min, double ep_max) {
    int ny = right - left, nz = end - start;
    double disc = psi.disc;
    int *map = malloc(sizeof(int) * (ny > 0 && nz > 0 ? (size_t)ny * nz : 1));
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nz; j++) {
            long long q = 3 * ((i + left) + (long long)psi.i * (y_slice + (long long)psi.j * (j + start)));
            map[i + j * ny] = psi.ep[q] > (ep_min + ep_max) / 2 ? 1 : 0;
        }
    FILE *f = fopen(filename, "wb");
    if (!f) {
        fprintf(stderr, "fdtd3d: cannot open %s\n", filename);
        free(map);
        return;
    }
    fwrite(&ny, sizeof(int), 1, f);
    fwrite(&nz, sizeof(int), 1, f);
    fwrite(&disc, sizeof(double), 1, f);
    if (ny > 0 && nz > 0) fwrite(map, sizeof(int), (size_t)ny * nz, f);
    fclose(f);
    free(map);
}
