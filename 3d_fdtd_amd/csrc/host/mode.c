/* Waveguide mode profiles (mode.h), host side.  Field layout: 6 complex values per transverse
 * cell (Ex,Ey,Ez,Hx,Hy,Hz), cell index i + j*mode.i (mode.c:4-35).  The inner products and the
 * in-place power normalisation keep the reference's summation order (mode.c:1237-1379): the
 * FDTD source/overlap calls renormalise the caller's mode on every call, and callers observe
 * the resulting values. */
#include "mode.h"

#include <string.h>

static complex cz(void) { return complex_zero(); }

mode mode_new(int i, int j, double disc) {
    mode m;
    m.i = i;
    m.j = j;
    m.disc = disc;
    int n = i * j;
    m.pml = matrix_new(2 * n, 1);
    m.ep = matrix_new(3 * n, 1);
    m.mu = matrix_new(3 * n, 1);
    m.fields = matrix_new(6 * n, 1);
    for (int q = 0; q < 2 * n; q++) m.pml.m[q] = complex_from_doubles(1, 0);
    return m;
}

void mode_free(mode psi) {
    matrix_free(psi.pml);
    matrix_free(psi.ep);
    matrix_free(psi.mu);
    matrix_free(psi.fields);
}

mode mode_copy(mode psi) {
    mode c = mode_new(psi.i, psi.j, psi.disc);
    int n = psi.i * psi.j;
    memcpy(c.pml.m, psi.pml.m, sizeof(complex) * 2 * n);
    memcpy(c.ep.m, psi.ep.m, sizeof(complex) * 3 * n);
    memcpy(c.mu.m, psi.mu.m, sizeof(complex) * 3 * n);
    memcpy(c.fields.m, psi.fields.m, sizeof(complex) * 6 * n);
    return c;
}

static void same_shape(mode a, mode b, int check_disc) {
    if (a.i != b.i || a.j != b.j || (check_disc && a.disc != b.disc)) {
        printf("Mode dimensions do not match for addition!\n");
        abort();
    }
}

void mode_copy_fields(mode psi, mode phi) {
    same_shape(psi, phi, 1);
    memcpy(psi.fields.m, phi.fields.m, sizeof(complex) * 6 * psi.i * psi.j);
}

void mode_smul(complex scalar, mode psi) {
    for (int q = 0; q < 6 * psi.i * psi.j; q++) psi.fields.m[q] = complex_mul(scalar, psi.fields.m[q]);
}

mode mode_linear_combination(mode psi, mode phi, complex s) {
    same_shape(psi, phi, 0);
    mode r = mode_copy(psi);
    for (int q = 0; q < 6 * psi.i * psi.j; q++)
        r.fields.m[q] = complex_add(psi.fields.m[q], complex_mul(s, phi.fields.m[q]));
    return r;
}

mode mode_linear_combination_3(mode psi, mode phi, mode alpha, complex s_phi, complex s_alpha) {
    same_shape(psi, phi, 0);
    mode r = mode_copy(psi);
    for (int q = 0; q < 6 * psi.i * psi.j; q++)
        r.fields.m[q] = complex_add(complex_add(psi.fields.m[q], complex_mul(s_phi, phi.fields.m[q])),
                                    complex_mul(s_alpha, alpha.fields.m[q]));
    return r;
}

/* Gaussian transverse H (Hy for polarization 0, Hx for 1) of width sigma cells centred on the
 * window, completed to E and Hz by the discrete curl/divergence relations at beta = omega and
 * power-normalised (mode.c:174-234). */
mode mode_gaussian(int i, int j, double disc, complex ep, complex mu, double omega, double sigma, int polarization) {
    mode g = mode_new(i, j, disc);
    int n = i * j;
    double xc = (double)i / 2, yc = (double)j / 2;
    for (int q = 0; q < 3 * n; q++) {
        g.ep.m[q] = ep;
        g.mu.m[q] = mu;
    }
    for (int q = 0; q < n; q++) {
        double x = (double)(q % i), y = (double)q / i;
        double dxx = x - xc, dyy = y - yc;
        double r2 = dxx * dxx + dyy * dyy;
        complex v = complex_from_doubles(exp(-r2 / (2 * sigma * sigma)), 0);
        g.fields.m[6 * q + (polarization ? 3 : 4)] = v;
    }
    mode_apply_H_curl(g, complex_from_doubles(omega, 0), omega);
    mode_apply_H_div(g, complex_from_doubles(omega, 0));
    mode_power_normalize(g);
    return g;
}

void mode_chop_bottom_half(mode psi) {
    for (int ii = 0; ii < psi.i; ii++)
        for (int jj = 0; jj < psi.j / 2; jj++)
            for (int m = 0; m < 6; m++) psi.fields.m[6 * (ii + jj * psi.i) + m] = cz();
}

/* ---- plots: text dumps into $FDTD3D_PLOT_DIR ------------------------------------------- */

static FILE *plot_file(const char *name) {
    const char *dir = getenv("FDTD3D_PLOT_DIR");
    if (!dir || !*dir) return NULL;
    char path[4096];
    snprintf(path, sizeof path, "%s/%s", dir, name);
    return fopen(path, "w");
}

static void plot_component(mode psi, const char *name, int comp, int real_part) {
    FILE *f = plot_file(name);
    if (!f) return;
    for (int jj = 0; jj < psi.j; jj++) {
        for (int ii = 0; ii < psi.i; ii++) {
            complex v = psi.fields.m[6 * (ii + jj * psi.i) + comp];
            fprintf(f, "%g\t", real_part ? v.re : complex_mag(v));
        }
        fprintf(f, "\n");
    }
    fclose(f);
}

void mode_plot(mode psi) {
    static const char *names[6] = {"mode_Ex.txt", "mode_Ey.txt", "mode_Ez.txt",
                                   "mode_Hx.txt", "mode_Hy.txt", "mode_Hz.txt"};
    for (int c = 0; c < 6; c++) plot_component(psi, names[c], c, 0);
}

void mode_plot_real(mode psi) {
    static const char *names[6] = {"mode_re_Ex.txt", "mode_re_Ey.txt", "mode_re_Ez.txt",
                                   "mode_re_Hx.txt", "mode_re_Hy.txt", "mode_re_Hz.txt"};
    for (int c = 0; c < 6; c++) plot_component(psi, names[c], c, 1);
}

void mode_plot_ex(mode psi) { plot_component(psi, "mode_Ex.txt", 0, 0); }
void mode_plot_hy(mode psi) { plot_component(psi, "mode_Hy.txt", 4, 0); }

/* ---- text file format (mode.c:843-1038) ------------------------------------------------- */

static void print_cell(FILE *f, mode psi, int ii, int jj) {
    int e = 3 * (ii + jj * psi.i), h = 6 * (ii + jj * psi.i);
    for (int k = 0; k < 3; k++) fprintf(f, "%f\t%f\t", psi.ep.m[e + k].re, psi.ep.m[e + k].im);
    for (int k = 0; k < 3; k++) fprintf(f, "%f\t%f\t", psi.mu.m[e + k].re, psi.mu.m[e + k].im);
    for (int k = 0; k < 6; k++)
        fprintf(f, k < 5 ? "%f\t%f\t" : "%f\t%f\n", psi.fields.m[h + k].re, psi.fields.m[h + k].im);
}

void mode_file_print(FILE *f, mode psi, int dimension) {
    fprintf(f, "%d\t%d\t%f\t%d\n", psi.i, psi.j, psi.disc, dimension);
    if (dimension == 2) {
        for (int ii = 0; ii < psi.i; ii++)
            for (int jj = 0; jj < psi.j; jj++) print_cell(f, psi, ii, jj);
    } else if (dimension == 1) {
        for (int jj = 0; jj < psi.j; jj++) print_cell(f, psi, psi.i / 2, jj);
    }
}

static int read_pair(FILE *f, complex *out) {
    double re, im;
    if (fscanf(f, "%lf", &re) != 1 || fscanf(f, "%lf", &im) != 1) return 0;
    *out = complex_from_doubles(re, im);
    return 1;
}

/* Header "i j disc dimension"; dimension 2 lists every cell (i outer, j inner), dimension 1
 * one row per j that is replicated along i.  A missing or truncated file aborts with a
 * message (the reference reads garbage there, mode.c:949-1038). */
mode mode_file_load(FILE *f) {
    int x, y, dimension;
    double disc;
    if (!f) {
        fprintf(stderr, "mode_file_load: no file\n");
        abort();
    }
    if (fscanf(f, "%d", &x) != 1 || fscanf(f, "%d", &y) != 1 || fscanf(f, "%lf", &disc) != 1 ||
        fscanf(f, "%d", &dimension) != 1 || x <= 0 || y <= 0) {
        fprintf(stderr, "mode_file_load: bad header\n");
        abort();
    }
    mode m = mode_new(x, y, disc);
    int ok = 1;
    if (dimension == 1) {
        for (int jj = 0; jj < y && ok; jj++) {
            complex v[12];
            for (int q = 0; q < 12 && ok; q++) ok = read_pair(f, &v[q]);
            for (int ii = 0; ii < x && ok; ii++) {
                int e = 3 * (ii + jj * x), h = 6 * (ii + jj * x);
                for (int k = 0; k < 3; k++) {
                    m.ep.m[e + k] = v[k];
                    m.mu.m[e + k] = v[3 + k];
                }
                for (int k = 0; k < 6; k++) m.fields.m[h + k] = v[6 + k];
            }
        }
    } else if (dimension == 2) {
        for (int ii = 0; ii < x && ok; ii++)
            for (int jj = 0; jj < y && ok; jj++) {
                int e = 3 * (ii + jj * x), h = 6 * (ii + jj * x);
                for (int k = 0; k < 3 && ok; k++) ok = read_pair(f, &m.ep.m[e + k]);
                for (int k = 0; k < 3 && ok; k++) ok = read_pair(f, &m.mu.m[e + k]);
                for (int k = 0; k < 6 && ok; k++) ok = read_pair(f, &m.fields.m[h + k]);
            }
    }
    if (!ok) {
        fprintf(stderr, "mode_file_load: truncated file\n");
        abort();
    }
    return m;
}

double mode_max_index(mode psi) {
    double best = 0;
    for (int q = 0; q < psi.i * psi.j; q++) {
        complex n = complex_sqrt(psi.ep.m[q]);
        if (n.re > best) best = n.re;
    }
    return best;
}

void mode_initialize_ones(mode psi) {
    for (int q = 0; q < 6 * psi.i * psi.j; q++) psi.fields.m[q] = complex_from_doubles(1, 0);
}

/* ---- inner products and normalisation (mode.c:1208-1379) -------------------------------- */

/* disc^2 sum over every stored value of conj(psi) phi, in storage order (all six components). */
complex mode_inner_product(mode psi, mode phi) {
    same_shape(psi, phi, 1);
    complex s = cz();
    for (int q = 0; q < 6 * psi.i * psi.j; q++)
        s = complex_add(s, complex_mul(complex_conj(psi.fields.m[q]), phi.fields.m[q]));
    return complex_smul(psi.disc * psi.disc, s);
}

complex mode_H_inner_product(mode psi, mode phi) {
    same_shape(psi, phi, 1);
    complex s = cz();
    for (int q = 0; q < psi.i * psi.j; q++) {
        s = complex_add(s, complex_mul(complex_conj(psi.fields.m[6 * q + 3]), phi.fields.m[6 * q + 3]));
        s = complex_add(s, complex_mul(complex_conj(psi.fields.m[6 * q + 4]), phi.fields.m[6 * q + 4]));
    }
    return complex_smul(psi.disc * psi.disc, s);
}

/* The complex helpers of complex.c, inlined for the per-cell loops below: the same double
 * operations in the same order (so the same bits), without a call per operation; every drop-in
 * source and overlap call power-normalises its mode, so these loops run several times a step. */
static inline complex c_mul(complex x, complex y) {
    complex z = {x.re * y.re - x.im * y.im, x.re * y.im + x.im * y.re};
    return z;
}
static inline complex c_conj(complex x) {
    complex z = {x.re, -x.im};
    return z;
}
static inline complex c_add(complex x, complex y) {
    complex z = {x.re + y.re, x.im + y.im};
    return z;
}
static inline complex c_neg0(complex x) { /* 0 - x, as complex_sub(complex_zero(), x) */
    complex z = {0.0 - x.re, 0.0 - x.im};
    return z;
}

/* -dx^2/2 sum conj(Hx_psi)Ey_phi - conj(Hy_psi)Ex_phi + Hx_phi conj(Ey_psi) - Hy_phi conj(Ex_psi),
 * i outer and j inner. */
complex mode_power_inner_product(mode psi, mode phi) {
    same_shape(psi, phi, 1);
    complex s = cz();
    for (int ii = 0; ii < psi.i; ii++)
        for (int jj = 0; jj < psi.j; jj++) {
            const complex *a = psi.fields.m + 6 * (ii + jj * psi.i);
            const complex *b = phi.fields.m + 6 * (ii + jj * psi.i);
            complex p1 = c_mul(c_conj(a[3]), b[1]);
            complex p2 = c_neg0(c_mul(c_conj(a[4]), b[0]));
            complex p3 = c_mul(b[3], c_conj(a[1]));
            complex p4 = c_neg0(c_mul(b[4], c_conj(a[0])));
            s = c_add(s, c_add(c_add(p1, p2), c_add(p3, p4)));
        }
    return complex_smul(-0.5 * psi.disc * psi.disc, s);
}

double mode_norm(mode psi) { return sqrt(complex_mag(mode_inner_product(psi, psi))); }
double mode_H_norm(mode psi) { return sqrt(complex_mag(mode_H_inner_product(psi, psi))); }
double mode_power_norm(mode psi) { return sqrt(complex_mag(mode_power_inner_product(psi, psi))); }

static void scale_fields(mode psi, double norm) {
    double inv = 1.0 / norm;
    /* x * 1.0 == x for every double, so a mode already normalised (the state of every call after
     * the first in a driver's loop) is left as the pass would leave it, without the pass */
    if (inv == 1.0) return;
    for (int q = 0; q < 6 * psi.i * psi.j; q++) {
        psi.fields.m[q].re = inv * psi.fields.m[q].re;
        psi.fields.m[q].im = inv * psi.fields.m[q].im;
    }
}

void mode_normalize(mode psi) { scale_fields(psi, mode_norm(psi)); }
void mode_H_normalize(mode psi) { scale_fields(psi, mode_H_norm(psi)); }
void mode_power_normalize(mode psi) { scale_fields(psi, mode_power_norm(psi)); }

/* ---- cross-section setters (mode.c:2199-2328) ------------------------------------------- */

static void set_cell(mode psi, int ii, int jj, complex ep, complex mu) {
    int e = 3 * (ii + jj * psi.i);
    for (int k = 0; k < 3; k++) {
        psi.ep.m[e + k] = ep;
        psi.mu.m[e + k] = mu;
    }
}

void mode_set_background(mode psi, complex ep, complex mu) {
    for (int ii = 0; ii < psi.i; ii++)
        for (int jj = 0; jj < psi.j; jj++) set_cell(psi, ii, jj, ep, mu);
}

void mode_set_bottom_half_plane(mode psi, complex ep, complex mu, int top) {
    for (int ii = 0; ii < psi.i; ii++)
        for (int jj = 0; jj < psi.j; jj++)
            if (jj < top) set_cell(psi, ii, jj, ep, mu);
}

void mode_set_top_half_plane(mode psi, complex ep, complex mu, int bottom) {
    for (int ii = 0; ii < psi.i; ii++)
        for (int jj = 0; jj < psi.j; jj++)
            if (jj >= bottom) set_cell(psi, ii, jj, ep, mu);
}

void mode_set_left_half_plane(mode psi, complex ep, complex mu, int right) {
    for (int ii = 0; ii < psi.i; ii++)
        for (int jj = 0; jj < psi.j; jj++)
            if (ii < right) set_cell(psi, ii, jj, ep, mu);
}

void mode_set_right_half_plane(mode psi, complex ep, complex mu, int left) {
    for (int ii = 0; ii < psi.i; ii++)
        for (int jj = 0; jj < psi.j; jj++)
            if (ii >= left) set_cell(psi, ii, jj, ep, mu);
}

void mode_make_rectangle(mode psi, complex ep, complex mu, int left, int right, int bottom, int top) {
    for (int ii = 0; ii < psi.i; ii++)
        for (int jj = 0; jj < psi.j; jj++)
            if (ii >= left && ii < right && jj >= bottom && jj < top) set_cell(psi, ii, jj, ep, mu);
}

/* ---- discrete curl / divergence for the transverse H (mode.c:2330-2397, 3350-3402) ------- */

/* E = (1/(i omega ep)) curl H with z dependence e^{i beta z}; backward differences in x, y
 * stretched by the PML factors sx, sy; interior cells only. */
void mode_apply_H_curl(mode psi, complex beta, double omega) {
    double inv_d = 1 / (psi.disc);
    complex mi_over_w = complex_smul(1 / omega, complex_from_doubles(0, -1));
    complex ib = complex_mul(complex_from_doubles(0, 1), beta);
    for (int ii = 1; ii < psi.i - 1; ii++)
        for (int jj = 1; jj < psi.j - 1; jj++) {
            int e = 3 * (ii + jj * psi.i), p = 2 * (ii + jj * psi.i);
            complex *F = psi.fields.m + 2 * e;
            const complex *L = F - 6, *D = F - 6 * psi.i;
            complex sx = psi.pml.m[p], sy = psi.pml.m[p + 1];
            complex hx = F[3], hy = F[4], hz = F[5];
            complex dx_hy = complex_div(complex_smul(inv_d, complex_sub(hy, L[4])), sx);
            complex dx_hz = complex_div(complex_smul(inv_d, complex_sub(hz, L[5])), sx);
            complex dy_hx = complex_div(complex_smul(inv_d, complex_sub(hx, D[3])), sy);
            complex dy_hz = complex_div(complex_smul(inv_d, complex_sub(hz, D[5])), sy);
            complex iex = complex_reciprocal(psi.ep.m[e]);
            complex iey = complex_reciprocal(psi.ep.m[e + 1]);
            complex iez = complex_reciprocal(psi.ep.m[e + 2]);
            F[0] = complex_mul(mi_over_w, complex_mul(iex, complex_add(dy_hz, complex_mul(ib, hy))));
            F[1] = complex_mul(mi_over_w, complex_smul(-1, complex_mul(iey, complex_add(dx_hz, complex_mul(ib, hx)))));
            F[2] = complex_mul(mi_over_w, complex_mul(iez, complex_sub(dx_hy, dy_hx)));
        }
}

/* Hz = -(i/beta) (d(mu_x Hx)/dx + d(mu_y Hy)/dy), forward differences, stretched. */
void mode_apply_H_div(mode psi, complex beta) {
    double inv_d = 1 / (psi.disc);
    complex mi_over_b = complex_div(complex_from_doubles(0, -1), beta);
    for (int ii = 1; ii < psi.i - 1; ii++)
        for (int jj = 1; jj < psi.j - 1; jj++) {
            int e = 3 * (ii + jj * psi.i), p = 2 * (ii + jj * psi.i);
            int e_r = e + 3, e_u = e + 3 * psi.i;
            complex *F = psi.fields.m + 2 * e;
            complex bx = complex_mul(psi.mu.m[e], F[3]);
            complex by = complex_mul(psi.mu.m[e + 1], F[4]);
            complex bx_r = complex_mul(psi.mu.m[e_r], F[6 + 3]);
            complex by_u = complex_mul(psi.mu.m[e_u + 1], F[6 * psi.i + 4]);
            complex sx = psi.pml.m[p], sy = psi.pml.m[p + 1];
            complex dbx = complex_div(complex_smul(inv_d, complex_sub(bx_r, bx)), sx);
            complex dby = complex_div(complex_smul(inv_d, complex_sub(by_u, by)), sy);
            F[5] = complex_mul(mi_over_b, complex_add(dbx, dby));
        }
}
