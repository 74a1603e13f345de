/*
 * Finite-difference waveguide mode solver of the drop-in surface (mode.c:1164-1633, 1634-1791,
 * 2796-2943, 3461-3616, 3789-3826, 4494-4526): the source / overlap input of the FDTD runs.
 *
 * The discrete problem is the reference's: on an i x j transverse grid (stretch factors sx, sy for
 * a complex-coordinate PML), the transverse H field of a mode with propagation constant beta
 * satisfies  Op(Hx, Hy) = beta^2 (Hx, Hy)  with
 *   Op_x = d_x^-( D ) + eps_y d_y^+( C ) + w^2 eps_y mu_x Hx
 *   Op_y = d_y^-( D ) - eps_x d_x^+( C ) + w^2 eps_x mu_y Hy
 *   D = (d_x^+(mu_x Hx) + d_y^+(mu_y Hy)) / mu_z,   C = (d_y^- Hx - d_x^- Hy) / eps_z,
 * every difference divided by disc and by the stretch factor of the cell it is taken at
 * (mode.c:2860-2935).  The solver is a restarted Arnoldi iteration in the H inner product
 * (mode.c:1237): each restart builds a `size`-dimensional Krylov space from the current vector,
 * takes the Ritz pair whose beta^2 has the largest real part and restarts from its Ritz vector,
 * until |n_eff| changes by at most tol between restarts (mode.c:3485-3596); Hz then follows from
 * div B = 0 and E from the curl of H (mode.c:3598-3603).
 *
 * Differences from the reference, none of which changes the discrete problem:
 *   - the small Hessenberg eigenproblem is solved to working precision (shifted QR, inverse
 *     iteration) where the reference stops its QL iteration at an off-diagonal ratio of 1e-3
 *     (matrix.c:672-871), so beta here is the converged Ritz value;
 *   - the sub-diagonal entry after the last Krylov vector, which the reference writes one row past
 *     its size x size matrix (mode.c:3547), is not stored (it does not enter the Ritz pairs);
 *   - beta starts at 0 for the first curl (the reference reads it uninitialised; it only shapes
 *     the E components, which the iteration does not use);
 *   - no progress lines are printed, and at most kMaxRestarts restarts run.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "complex.h"
#include "matrix.h"
#include "mode.h"

enum { kMaxRestarts = 1000 };

/* ---- small complex helpers on the library's complex type ---- */

static inline complex cx(double re, double im) {
    complex z;
    z.re = re;
    z.im = im;
    return z;
}
static inline complex cadd(complex a, complex b) { return cx(a.re + b.re, a.im + b.im); }
static inline complex csub(complex a, complex b) { return cx(a.re - b.re, a.im - b.im); }
static inline complex cmul(complex a, complex b) { return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline complex cscale(double s, complex a) { return cx(s * a.re, s * a.im); }
static inline complex cconj(complex a) { return cx(a.re, -a.im); }
static inline double cabs2(complex a) { return a.re * a.re + a.im * a.im; }
static inline complex cdivz(complex a, complex b) {
    const double d = cabs2(b);
    return cx((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}

/* ---- initial vectors, boundary conditions, PML profile ---- */

void mode_initialize_general(mode psi) {
    /* Hx = Hy = 1 + 3 (x - x_c)/i + (y - y_c)/j (mode.c:1164-1190) */
    const int xc = psi.i / 2, yc = psi.j / 2;
    for (int c = 0; c < psi.i * psi.j; c++) {
        const double x = (double)(c % psi.i), y = (double)c / psi.i;
        const double v = 1 + ((x - xc) * 3) / psi.i + ((y - yc) * 1) / psi.j;
        psi.fields.m[6 * c + 3] = cx(v, 0);
        psi.fields.m[6 * c + 4] = cx(v, 0);
    }
}

void mode_initialize_rand(mode psi) {
    /* Hx, Hy of every cell from successive rand() values, in cell order (mode.c:1192-1206) */
    for (int c = 0; c < psi.i * psi.j; c++) {
        psi.fields.m[6 * c + 3] = cx((double)rand(), 0);
        psi.fields.m[6 * c + 4] = cx((double)rand(), 0);
    }
}

static void zero_cell(mode psi, int c) {
    for (int m = 0; m < 6; m++) psi.fields.m[6 * c + m] = cx(0, 0);
}

static void copy_cell(mode psi, int dst, int src) {
    for (int m = 0; m < 6; m++) psi.fields.m[6 * dst + m] = psi.fields.m[6 * src + m];
}

void mode_apply_zero_bc_horizontal(mode psi) {
    for (int y = 0; y < psi.j; y++) {
        zero_cell(psi, y * psi.i + psi.i - 1);
        zero_cell(psi, y * psi.i);
    }
}

void mode_apply_zero_bc_vertical(mode psi) {
    for (int x = 1; x < psi.i - 1; x++) {
        zero_cell(psi, x + (psi.j - 1) * psi.i);
        zero_cell(psi, x);
    }
}

void mode_apply_zero_bc(mode psi) {
    mode_apply_zero_bc_horizontal(psi);
    mode_apply_zero_bc_vertical(psi);
}

void mode_apply_zero_gradient_bc_horizontal(mode psi) {
    for (int y = 0; y < psi.j; y++) {
        copy_cell(psi, y * psi.i + psi.i - 1, y * psi.i + psi.i - 2);
        copy_cell(psi, y * psi.i, y * psi.i + 1);
    }
}

void mode_apply_zero_gradient_bc_vertical(mode psi) {
    for (int x = 1; x < psi.i - 1; x++) {
        copy_cell(psi, x + (psi.j - 1) * psi.i, x + (psi.j - 2) * psi.i);
        copy_cell(psi, x, x + psi.i);
    }
}

/* Zero-gradient copy on all four edges (mode.c:1484): the side columns, then the top and bottom
 * rows without the corners.  The reference's left-edge copy takes component c of the edge cell
 * from storage slot 6 + c + 1 after it (c = 1..5; c = 0 from slot 6), one slot further than the
 * right edge does; that shift is kept so results match, and a slot past the array's end (a mode
 * one or two cells wide) is not read. */
void mode_apply_zero_gradient_bc(mode psi) {
    const int total = 6 * psi.i * psi.j;
    for (int y = 0; y < psi.j; y++) {
        copy_cell(psi, y * psi.i + psi.i - 1, y * psi.i + psi.i - 2);
        const int left = 6 * y * psi.i;
        psi.fields.m[left] = psi.fields.m[left + 6];
        for (int c = 1; c < 6; c++)
            if (left + 7 + c < total) psi.fields.m[left + c] = psi.fields.m[left + 7 + c];
    }
    mode_apply_zero_gradient_bc_vertical(psi);
}

void mode_apply_zero_bc_pml(mode psi) {
    /* every cell where either stretch factor has a positive imaginary part (mode.c:1587-1612) */
    for (int c = 0; c < psi.i * psi.j; c++)
        if (psi.pml.m[2 * c].im > 1e-9 || psi.pml.m[2 * c + 1].im > 1e-9) zero_cell(psi, c);
}

void mode_apply_bc(mode psi, int confinement) {
    /* confinement 1: slab (zero top/bottom, zero-gradient sides); 2: zero on every edge */
    if (confinement != 1 && confinement != 2) {
        printf("Mode confinement dimension is not 1 or 2!\n");
        abort();
    }
    if (confinement == 1) {
        mode_apply_zero_bc_vertical(psi);
        mode_apply_zero_gradient_bc_horizontal(psi);
    } else {
        mode_apply_zero_bc(psi);
    }
}

void mode_set_pml(mode psi, double max_ep_im, int order, int pml_thickness) {
    /* s = 1 + i a d^order / T^order over the outer T cells of each edge, d the depth counted from
     * T at the outer cell (lo side) and from 1 at the first PML cell (hi side); order 0 is a
     * constant a (mode.c:1634-1790) */
    if (order < 0 || order > 4) {
        printf("Order not valid for set pml!\n");
        abort();
    }
    double norm = 1.0;
    for (int q = 0; q < order; q++) norm *= pml_thickness;
    const double step = max_ep_im / norm;
    for (int y = 0; y < psi.j; y++)
        for (int x = 0; x < psi.i; x++) {
            const int p = 2 * (x + y * psi.i);
            int d;
            if (x < pml_thickness) {
                d = pml_thickness - x;
                psi.pml.m[p] = cx(1, step * pow((double)d, order));
            }
            if (x >= psi.i - pml_thickness) {
                d = x - (psi.i - 1 - pml_thickness);
                psi.pml.m[p] = cx(1, step * pow((double)d, order));
            }
            if (y < pml_thickness) {
                d = pml_thickness - y;
                psi.pml.m[p + 1] = cx(1, step * pow((double)d, order));
            }
            if (y >= psi.j - pml_thickness) {
                d = y - (psi.j - 1 - pml_thickness);
                psi.pml.m[p + 1] = cx(1, step * pow((double)d, order));
            }
        }
}

void mode_apply_lateral_symmetry(mode psi) {
    /* average with the mirror image x -> i-1-x (Hx odd, Hy even), choosing the sign that keeps
     * the larger part (mode.c:4494-4526) */
    mode r = mode_new(psi.i, psi.j, psi.disc);
    for (int y = 0; y < psi.j; y++)
        for (int x = 0; x < psi.i; x++) {
            const int c = x + y * psi.i, s = (psi.i - 1 - x) + y * psi.i;
            r.fields.m[6 * s + 3] = csub(cx(0, 0), psi.fields.m[6 * c + 3]);
            r.fields.m[6 * s + 4] = psi.fields.m[6 * c + 4];
        }
    const complex ip = mode_H_inner_product(psi, r);
    mode sym = mode_linear_combination(psi, r, cx(ip.re > 0 ? 1.0 : -1.0, 0));
    mode_free(r);
    mode_copy_fields(psi, sym);
    mode_free(sym);
}

/* ---- the transverse-H operator ---- */

mode mode_apply_beta_squared_operator(mode psi, double omega) {
    const int I = psi.i, J = psi.j;
    const double h = 1.0 / psi.disc;
    mode out = mode_new(I, J, psi.disc);
    for (int q = 0; q < 3 * I * J; q++) {
        out.ep.m[q] = psi.ep.m[q];
        out.mu.m[q] = psi.mu.m[q];
    }
    for (int q = 0; q < 2 * I * J; q++) out.pml.m[q] = psi.pml.m[q];
    if (I < 3 || J < 3) return out;
    const complex *F = psi.fields.m, *ep = psi.ep.m, *mu = psi.mu.m, *s = psi.pml.m;
#define HX(x, y) F[6 * ((x) + (y) * I) + 3]
#define HY(x, y) F[6 * ((x) + (y) * I) + 4]
#define EP(x, y, m) ep[3 * ((x) + (y) * I) + (m)]
#define MU(x, y, m) mu[3 * ((x) + (y) * I) + (m)]
#define SX(x, y) s[2 * ((x) + (y) * I)]
#define SY(x, y) s[2 * ((x) + (y) * I) + 1]
    /* D on [0, I-2] x [0, J-2], C on [1, I-1] x [1, J-1] */
    complex *D = calloc((size_t)I * J, sizeof(complex)), *Cz = calloc((size_t)I * J, sizeof(complex));
    if (!D || !Cz) {
        printf("mode operator: out of memory\n");
        abort();
    }
    for (int y = 0; y < J - 1; y++)
        for (int x = 0; x < I - 1; x++) {
            const complex bx = cdivz(cscale(h, csub(cmul(MU(x + 1, y, 0), HX(x + 1, y)), cmul(MU(x, y, 0), HX(x, y)))),
                                     SX(x, y));
            const complex by = cdivz(cscale(h, csub(cmul(MU(x, y + 1, 1), HY(x, y + 1)), cmul(MU(x, y, 1), HY(x, y)))),
                                     SY(x, y));
            D[x + y * I] = cdivz(cadd(bx, by), MU(x, y, 2));
        }
    for (int y = 1; y < J; y++)
        for (int x = 1; x < I; x++) {
            const complex hxy = cdivz(cscale(h, csub(HX(x, y), HX(x, y - 1))), SY(x, y));
            const complex hyx = cdivz(cscale(h, csub(HY(x, y), HY(x - 1, y))), SX(x, y));
            Cz[x + y * I] = cdivz(csub(hxy, hyx), EP(x, y, 2));
        }
    const double w2 = omega * omega;
    for (int y = 1; y < J - 1; y++)
        for (int x = 1; x < I - 1; x++) {
            const int c = x + y * I;
            const complex ox = cadd(cadd(cdivz(cscale(h, csub(D[c], D[c - 1])), SX(x, y)),
                                         cdivz(cmul(cscale(h, EP(x, y, 1)), csub(Cz[c + I], Cz[c])), SY(x, y))),
                                    cscale(w2, cmul(cmul(EP(x, y, 1), MU(x, y, 0)), HX(x, y))));
            const complex oy = cadd(cadd(cdivz(cscale(h, csub(D[c], D[c - I])), SY(x, y)),
                                         cdivz(cmul(cscale(-h, EP(x, y, 0)), csub(Cz[c + 1], Cz[c])), SX(x, y))),
                                    cscale(w2, cmul(cmul(EP(x, y, 0), MU(x, y, 1)), HY(x, y))));
            out.fields.m[6 * c + 3] = ox;
            out.fields.m[6 * c + 4] = oy;
        }
#undef HX
#undef HY
#undef EP
#undef MU
#undef SX
#undef SY
    free(D);
    free(Cz);
    return out;
}

/* ---- the restarted Arnoldi solver ---- */

static void fields_axpy(mode y, complex a, mode x) {
    for (int q = 0; q < 6 * y.i * y.j; q++) y.fields.m[q] = cadd(y.fields.m[q], cmul(a, x.fields.m[q]));
}

complex mode_solve_beta_squared(mode *psi, double omega, double tol, mode *prev_modes, int prev_num, int size,
                                int confinement, int lateral_symmetry) {
    if (size < 1) {
        printf("Mode solver: Krylov dimension must be positive!\n");
        abort();
    }
    mode *kr = malloc((size_t)(size + 1) * sizeof(mode));
    complex *lam = malloc((size_t)size * sizeof(complex));
    if (!kr || !lam) {
        printf("Mode solver: out of memory\n");
        abort();
    }
    complex beta = cx(0, 0);
    mode_initialize_rand(*psi);
    mode_H_normalize(*psi);
    double n_mag = 100.0, n_diff = 100.0;
    for (int restart = 0; fabs(n_diff) > tol && restart < kMaxRestarts; restart++) {
        /* the start vector: orthogonal to the previous modes, in the boundary conditions */
        mode_H_normalize(*psi);
        for (int q = prev_num - 1; q >= 0; q--) {
            fields_axpy(*psi, csub(cx(0, 0), mode_H_inner_product(prev_modes[q], *psi)), prev_modes[q]);
            mode_H_normalize(*psi);
        }
        mode_apply_H_curl(*psi, beta, omega);
        mode_apply_bc(*psi, confinement);
        if (lateral_symmetry) mode_apply_lateral_symmetry(*psi);
        mode_apply_zero_bc_pml(*psi);
        /* Arnoldi: H(j, k) = <v_j, Op v_k>, modified Gram-Schmidt in the H inner product */
        matrix H = matrix_new(size, size);
        kr[0] = mode_copy(*psi);
        mode_H_normalize(kr[0]);
        mode_apply_bc(kr[0], confinement);
        for (int k = 1; k <= size; k++) {
            kr[k] = mode_apply_beta_squared_operator(kr[k - 1], omega);
            mode_apply_bc(kr[k], confinement);
            for (int j = 0; j < k; j++) {
                const complex hjk = mode_H_inner_product(kr[j], kr[k]);
                H.m[(k - 1) + j * size] = hjk;
                fields_axpy(kr[k], csub(cx(0, 0), hjk), kr[j]);
            }
            const double nrm = mode_H_norm(kr[k]);
            if (k < size) H.m[(k - 1) + k * size] = cx(nrm, 0);
            if (nrm > 0) mode_H_normalize(kr[k]);
        }
        /* the Ritz value with the largest real part (first such, as long as it is positive) */
        matrix Dg = matrix_diagonalize(H, 1e-3, 1000);
        int best = 0;
        double top = 0;
        for (int q = 0; q < size; q++) {
            lam[q] = Dg.m[q + q * size];
            if (lam[q].re > top) {
                top = lam[q].re;
                best = q;
            }
        }
        matrix_free(Dg);
        matrix y = matrix_eigenvector(H, lam[best]);
        for (int q = 0; q < 6 * psi->i * psi->j; q++) psi->fields.m[q] = cx(0, 0);
        for (int j = 0; j < size; j++) fields_axpy(*psi, y.m[j], kr[j]);
        matrix_free(y);
        matrix_free(H);
        for (int k = 0; k <= size; k++) mode_free(kr[k]);
        beta = complex_sqrt(lam[best]);
        const double n_old = n_mag;
        n_mag = sqrt(cabs2(beta)) / omega;
        n_diff = n_mag - n_old;
    }
    mode_apply_H_div(*psi, beta);
    mode_apply_bc(*psi, confinement);
    mode_apply_H_curl(*psi, beta, omega);
    mode_apply_bc(*psi, confinement);
    mode_H_normalize(*psi);
    free(kr);
    free(lam);
    return beta;
}

/* ---- SI units (mode.c:3789-3826) ---- */

complex mode_true_beta(complex beta) { return cscale(1e6, beta); }

mode mode_true_fields(mode psi) {
    const double ep0 = 8.85418782e-12, mu0 = 1.25663706e-6, z0 = sqrt(mu0 / ep0);
    mode t = mode_copy(psi);
    for (int c = 0; c < psi.i * psi.j; c++)
        for (int m = 0; m < 3; m++) t.fields.m[6 * c + m] = cscale(z0, psi.fields.m[6 * c + m]);
    mode_power_normalize(t);
    return t;
}
