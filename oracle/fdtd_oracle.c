/*
 * fdtd_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never the thing measured or shipped).
 *
 * Clean-room FP64 restatement of the reference FDTD hot path.  Reference-compatible routines
 * keep the reference's floating-point operation order so that they are bit-for-bit equal to
 * the compiled reference (oracle/_ref) on the fixtures in tests/golden/; each routine cites
 * the reference lines it follows.  Extensions (CPML, multi-frequency scalar DFT) have no
 * reference counterpart; their formulas are stated below and pinned by invariants in tests/
 * (zero-width CPML == reference step bit-for-bit; CPML reflection vs a large-domain run;
 * pure-tone DFT closed form).
 *
 * Pinning status per routine (tests that pin it):
 *   orc_complex_exp/angle/sqrt  -- fixture F9 (tests/test_oracle_golden.py) + libm closed form
 *   orc_step (E/H/Mur/snapshot) -- fixtures F1, F2, F3 bit-exact
 *   orc_dft_volume/y_slice      -- fixtures F3 (5 frequencies), F4 bit-exact
 *   orc_mode_power_*            -- fixture F5 bit-exact, |P|=1 invariant
 *   orc_mode_source(_adjoint)   -- fixture F4 bit-exact
 *   orc_mode_overlap            -- fixture F4 bit-exact
 *   orc_set_taper/taper_insert  -- fixture F6 bit-exact
 *   orc_taper_*                 -- fixture F6 bit-exact
 *   orc_*_cpml, orc_cpml_profile-- zero-width == orc_step interior; reflection invariant
 *   orc_dft_scalar              -- pure-tone closed form
 *   orc_dft_box                 -- single frequency over the whole grid == orc_dft_volume (F3-pinned)
 *   orc_overlap_adjoint_source, -- finite differences: the gradient they give matches
 *   orc_dft_gradient               (a(eps + d) - a(eps - d)) / 2d of the overlap monitor
 */
#include "fdtd_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* The reference keeps two different truncated values of pi (timefield3d.c:3, adjoint3d.c:3
 * use 3.141592652; complex.c:3 uses 3.14159265359).  They are part of its phase semantics. */
#define PI_FIELD 3.141592652
#define PI_COMPLEX 3.14159265359

static inline long cidx(const orc_grid *g, int i, int j, int k) {
    return (long)i + (long)g->I * ((long)j + (long)g->J * (long)k);
}

/* ------------------------------------------------------------------------------------------ */
/* complex helpers                                                                            */
/* ------------------------------------------------------------------------------------------ */

/* Power series exp(x) until the current term's magnitude drops below 1e-30 (complex.c:169-193). */
void orc_complex_exp(double xre, double xim, double out[2]) {
    double term_re = 1.0, term_im = 0.0;   /* magnitude of the most recent term */
    double acc_re = 0.0, acc_im = 0.0;
    double pw_re = 1.0, pw_im = 0.0;       /* x^n */
    double fact = 1.0;
    int n = 0;
    while (sqrt(term_re * term_re + term_im * term_im) >= 1e-30) {
        double inv = 1.0 / fact;
        term_re = inv * pw_re;
        term_im = inv * pw_im;
        acc_re = acc_re + inv * pw_re;
        acc_im = acc_im + inv * pw_im;
        n++;
        fact *= n;
        double nr = pw_re * xre - pw_im * xim;
        double ni = pw_re * xim + pw_im * xre;
        pw_re = nr;
        pw_im = ni;
    }
    out[0] = acc_re;
    out[1] = acc_im;
}

/* Angle in [0, 2pi) via atan of the ratio plus quadrant fix-up (complex.c:15-35). */
double orc_complex_angle(double re, double im) {
    double a = atan(im / re);
    if (re < 0) a += PI_COMPLEX;
    else if (im < 0) a += 2 * PI_COMPLEX;
    return a;
}

/* Principal square root through magnitude and half angle (complex.c:151-167). */
void orc_complex_sqrt(double re, double im, double out[2]) {
    double mag = sqrt(sqrt(re * re + im * im));
    double ang = orc_complex_angle(re, im);
    if (ang > PI_COMPLEX) ang -= 2 * PI_COMPLEX;
    ang /= 2;
    double e[2];
    orc_complex_exp(0.0, ang, e);
    out[0] = mag * e[0];
    out[1] = mag * e[1];
}

/* phase - trunc(phase/2pi~)*2*pi~ with the int product formed first (timefield3d.c:2857). */
double orc_reduce_phase(double phase) {
    int turns = (int)(phase / (2 * PI_FIELD));
    return phase - (turns * 2) * PI_FIELD;
}

/* ------------------------------------------------------------------------------------------ */
/* Yee step, reference mode                                                                   */
/* ------------------------------------------------------------------------------------------ */

/* Copy the 2-cell shells of all six faces, all six components (timefield3d.c:246-346). */
void orc_snapshot_shells(const orc_grid *g) {
    const int N[3] = {g->I, g->J, g->K};
    for (int a = 0; a < 3; a++) {
        int b = (a + 1) % 3, c = (a + 2) % 3;
        for (int side = 0; side < 2; side++) {
            for (int layer = 0; layer < 2; layer++) {
                int na = side == 0 ? layer : N[a] - 1 - layer;
                for (int u = 0; u < N[b]; u++) {
                    for (int v = 0; v < N[c]; v++) {
                        int x[3];
                        x[a] = na; x[b] = u; x[c] = v;
                        long o = 6 * cidx(g, x[0], x[1], x[2]);
                        for (int m = 0; m < 6; m++) g->old[o + m] = g->f[o + m];
                    }
                }
            }
        }
    }
}

/* E interior update, semi-implicit conductivity, backward differences of H
 * (timefield3d.c:1759-1837); cells i,j,k in [1, N-2]. */
void orc_update_E(const orc_grid *g, double dt) {
    const int I = g->I, J = g->J, K = g->K;
    const double inv_dx = 1 / g->disc;
    const long sx = 6, sy = 6L * I, sz = 6L * I * J;
#pragma omp parallel for schedule(static)
    for (int k = 1; k < K - 1; k++) {
        for (int j = 1; j < J - 1; j++) {
            for (int i = 1; i < I - 1; i++) {
                long c = cidx(g, i, j, k);
                const double *f = g->f + 6 * c;
                double hx = f[3], hy = f[4], hz = f[5];
                double curl[3];
                curl[0] = inv_dx * ((hz - f[5 - sy]) - (hy - f[4 - sz]));
                curl[1] = inv_dx * ((hx - f[3 - sz]) - (hz - f[5 - sx]));
                curl[2] = inv_dx * ((hy - f[4 - sx]) - (hx - f[3 - sy]));
                double *fo = g->f + 6 * c;
                for (int m = 0; m < 3; m++) {
                    double ep = g->ep[3 * c + m];
                    double s = g->sig[6 * c + m];
                    double jsrc = g->src[6 * c + m];
                    double inv_den = 1.0 / (ep / dt + 0.5 * s);
                    double keep = ep / dt - 0.5 * s;
                    fo[m] = inv_den * (keep * fo[m] + curl[m] - jsrc);
                }
            }
        }
    }
}

/* H interior update with forward differences of E and the magnetic conductivity entering both
 * the semi-implicit coefficients and M = src + sigma*H (timefield3d.c:1679-1757). */
void orc_update_H(const orc_grid *g, double dt) {
    const int I = g->I, J = g->J, K = g->K;
    const double inv_dx = 1 / g->disc;
    const long sx = 6, sy = 6L * I, sz = 6L * I * J;
#pragma omp parallel for schedule(static)
    for (int k = 1; k < K - 1; k++) {
        for (int j = 1; j < J - 1; j++) {
            for (int i = 1; i < I - 1; i++) {
                long c = cidx(g, i, j, k);
                const double *f = g->f + 6 * c;
                double ex = f[0], ey = f[1], ez = f[2];
                double curl[3];
                curl[0] = inv_dx * ((f[2 + sy] - ez) - (f[1 + sz] - ey));
                curl[1] = inv_dx * ((f[0 + sz] - ex) - (f[2 + sx] - ez));
                curl[2] = inv_dx * ((f[1 + sx] - ey) - (f[0 + sy] - ex));
                double *fo = g->f + 6 * c;
                for (int m = 0; m < 3; m++) {
                    double h = fo[3 + m];
                    double msrc = g->src[6 * c + 3 + m] + h * g->sig[6 * c + 3 + m];
                    double mu = g->mu[3 * c + m];
                    double s = g->sig[6 * c + 3 + m];
                    double inv_den = 1.0 / (mu / dt + 0.5 * s);
                    double keep = mu / dt - 0.5 * s;
                    fo[3 + m] = inv_den * (keep * h - curl[m] - msrc);
                }
            }
        }
    }
}

/* First-order Mur on one face: F_b <- F_in^old + s*(F_in - F_b^old), s = (v dt - dx)/(v dt + dx),
 * v = 1/sqrt(ep*mu) of the boundary cell with the component pairing of
 * timefield3d.c:2275-2721 (E comp t: ep_t*mu_d; H comp t: ep_d*mu_t; d = the third axis). */
static void mur_face(const orc_grid *g, double dt, int field, int a, int side) {
    const int N[3] = {g->I, g->J, g->K};
    const int b = (a + 1) % 3, c = (a + 2) % 3;
    const int nb = side == 0 ? 0 : N[a] - 1;
    const int ni = side == 0 ? 1 : N[a] - 2;
    for (int u = 0; u < N[b]; u++) {
        for (int v = 0; v < N[c]; v++) {
            int xb[3], xi[3];
            xb[a] = nb; xb[b] = u; xb[c] = v;
            xi[a] = ni; xi[b] = u; xi[c] = v;
            long cb = cidx(g, xb[0], xb[1], xb[2]);
            long ci = cidx(g, xi[0], xi[1], xi[2]);
            const int tang[2] = {b, c};
            for (int q = 0; q < 2; q++) {
                int t = tang[q];
                int d = 3 - a - t;
                double e, m;
                if (field == 0) { e = g->ep[3 * cb + t]; m = g->mu[3 * cb + d]; }
                else            { e = g->ep[3 * cb + d]; m = g->mu[3 * cb + t]; }
                int comp = 3 * field + t;
                double speed = 1 / sqrt(e * m);
                double s = (speed * dt - g->disc) / (speed * dt + g->disc);
                double b_old = g->old[6 * cb + comp];
                double in_now = g->f[6 * ci + comp];
                double in_old = g->old[6 * ci + comp];
                g->f[6 * cb + comp] = in_old + s * (in_now - b_old);
            }
        }
    }
}

/* Face order x0, y0, z0, x_end, y_end, z_end: later faces win on edges and corners
 * (timefield3d.c:2499-2721 for E, 2275-2497 for H). */
void orc_mur_E(const orc_grid *g, double dt) {
    for (int side = 0; side < 2; side++)
        for (int a = 0; a < 3; a++) mur_face(g, dt, 0, a, side);
}

void orc_mur_H(const orc_grid *g, double dt) {
    for (int side = 0; side < 2; side++)
        for (int a = 0; a < 3; a++) mur_face(g, dt, 1, a, side);
}

/* One full step in the order of timefield3d_apply_timestep_omp_full (timefield3d.c:2787-2849). */
void orc_step(const orc_grid *g, double dt) {
    orc_snapshot_shells(g);
    orc_update_E(g, dt);
    orc_mur_E(g, dt);
    orc_update_H(g, dt);
    orc_mur_H(g, dt);
}

/* ------------------------------------------------------------------------------------------ */
/* Running DFT, reference semantics (timefield3d.c:2851-2975)                                 */
/* ------------------------------------------------------------------------------------------ */

static void dft_cells(const orc_grid *g, double tw_re, double tw_im, long c, double *F) {
    for (int m = 0; m < 6; m++) {
        double v = g->f[6 * c + m];
        F[2 * (6 * c + m)] = F[2 * (6 * c + m)] + v * tw_re;
        F[2 * (6 * c + m) + 1] = F[2 * (6 * c + m) + 1] + v * tw_im;
    }
}

void orc_dft_volume(const orc_grid *g, double phase, double *F) {
    double tw[2];
    orc_complex_exp(0.0, orc_reduce_phase(phase), tw);
    long n = (long)g->I * g->J * g->K;
#pragma omp parallel for schedule(static)
    for (long c = 0; c < n; c++) dft_cells(g, tw[0], tw[1], c, F);
}

void orc_dft_y_slice(const orc_grid *g, double phase, int j, double *F) {
    double tw[2];
    orc_complex_exp(0.0, orc_reduce_phase(phase), tw);
    for (int k = 0; k < g->K; k++)
        for (int i = 0; i < g->I; i++) dft_cells(g, tw[0], tw[1], cidx(g, i, j, k), F);
}

/* ------------------------------------------------------------------------------------------ */
/* Modes (mode.c:1267-1379) and plane sources / monitors (timefield3d.c:2977-3068, 3176-3239)  */
/* ------------------------------------------------------------------------------------------ */

/* -dx^2/2 * sum conj(Hx_psi)Ey_phi - conj(Hy_psi)Ex_phi + Hx_phi conj(Ey_psi) - Hy_phi conj(Ex_psi),
 * i outer, j inner (mode.c:1267-1310). */
void orc_mode_power_inner(int mi, int mj, double mdisc, const double *psi, const double *phi, double out[2]) {
    double sr = 0.0, si = 0.0;
    for (int i = 0; i < mi; i++) {
        for (int j = 0; j < mj; j++) {
            long o = 2L * 6 * (i + (long)j * mi);
            double pEx_r = psi[o + 0], pEx_i = psi[o + 1], pEy_r = psi[o + 2], pEy_i = psi[o + 3];
            double pHx_r = psi[o + 6], pHx_i = psi[o + 7], pHy_r = psi[o + 8], pHy_i = psi[o + 9];
            double qEx_r = phi[o + 0], qEx_i = phi[o + 1], qEy_r = phi[o + 2], qEy_i = phi[o + 3];
            double qHx_r = phi[o + 6], qHx_i = phi[o + 7], qHy_r = phi[o + 8], qHy_i = phi[o + 9];
            /* conj(a)*b */
            double p1r = pHx_r * qEy_r - (-pHx_i) * qEy_i, p1i = pHx_r * qEy_i + (-pHx_i) * qEy_r;
            double t2r = pHy_r * qEx_r - (-pHy_i) * qEx_i, t2i = pHy_r * qEx_i + (-pHy_i) * qEx_r;
            double p2r = 0.0 - t2r, p2i = 0.0 - t2i;
            /* b*conj(a) */
            double p3r = qHx_r * pEy_r - qHx_i * (-pEy_i), p3i = qHx_r * (-pEy_i) + qHx_i * pEy_r;
            double t4r = qHy_r * pEx_r - qHy_i * (-pEx_i), t4i = qHy_r * (-pEx_i) + qHy_i * pEx_r;
            double p4r = 0.0 - t4r, p4i = 0.0 - t4i;
            sr = sr + ((p1r + p2r) + (p3r + p4r));
            si = si + ((p1i + p2i) + (p3i + p4i));
        }
    }
    double s = -0.5 * mdisc * mdisc;
    out[0] = s * sr;
    out[1] = s * si;
}

/* Scale all components by 1/sqrt(|P|) (mode.c:1330-1379). */
void orc_mode_power_normalize(int mi, int mj, double mdisc, double *mode) {
    double p[2];
    orc_mode_power_inner(mi, mj, mdisc, mode, mode, p);
    double norm = sqrt(sqrt(p[0] * p[0] + p[1] * p[1]));
    double inv = 1.0 / norm;
    long n = 2L * 6 * mi * mj;
    for (long q = 0; q < n; q++) mode[q] = inv * mode[q];
}

/* Mode-window index of grid coordinate (int truncation toward zero, timefield3d.c:2997-3004). */
static inline int mode_index_1d(int n, int center, double ratio, int mn) {
    double dist = (n - center) * ratio;
    return (int)dist + mn / 2;
}

/* Soft mode source on plane k = cz.  E: J = A*ep*Re(mode e^{-i phase}); H: M = -(2r-1)*A*mu*(...),
 * stored as float; cells outside the mode window keep their previous value
 * (timefield3d.c:2977-3026). */
void orc_mode_source(const orc_grid *g, int cx, int cy, int cz, int mi, int mj, double mdisc, double *mode,
                     double phase, double amplitude, int reverse) {
    int msign = -1 + 2 * reverse;
    orc_mode_power_normalize(mi, mj, mdisc, mode);
    double ratio = g->disc / mdisc;
    for (int i = 0; i < g->I; i++) {
        for (int j = 0; j < g->J; j++) {
            int mx = mode_index_1d(i, cx, ratio, mi);
            int my = mode_index_1d(j, cy, ratio, mj);
            if (mx < 0 || mx >= mi || my < 0 || my >= mj) continue;
            long c = cidx(g, i, j, cz);
            for (int m = 0; m < 6; m++) {
                long mo = 2L * (6 * (mx + (long)mi * my) + m);
                double proj = mode[mo] * cos(phase) + mode[mo + 1] * sin(phase);
                if (m < 3) {
                    double ep = g->ep[3 * c + m];
                    g->src[6 * c + m] = (float)(amplitude * ep * proj);
                } else {
                    double mu = g->mu[3 * c + m % 3];
                    g->src[6 * c + m] = (float)((-msign) * amplitude * mu * proj);
                }
            }
        }
    }
}

/* CW adjoint source: J = (Re Hy, -Re Hx, 0), M = (-(2r-1) Re Ey, (2r-1) Re Ex, 0) of the mode times
 * e^{-i phase}; no material factor (timefield3d.c:3028-3068). */
void orc_mode_source_adjoint(const orc_grid *g, int cx, int cy, int cz, int mi, int mj, double mdisc, double *mode,
                             double phase, int reverse) {
    int msign = -1 + 2 * reverse;
    orc_mode_power_normalize(mi, mj, mdisc, mode);
    double ratio = g->disc / mdisc;
    for (int i = 0; i < g->I; i++) {
        for (int j = 0; j < g->J; j++) {
            int mx = mode_index_1d(i, cx, ratio, mi);
            int my = mode_index_1d(j, cy, ratio, mj);
            if (mx < 0 || mx >= mi || my < 0 || my >= mj) continue;
            long c = cidx(g, i, j, cz);
            long mo = 2L * 6 * (mx + (long)mi * my);
            double cp = cos(phase), sp = sin(phase);
            double rEx = mode[mo + 0] * cp + mode[mo + 1] * sp;
            double rEy = mode[mo + 2] * cp + mode[mo + 3] * sp;
            double rHx = mode[mo + 6] * cp + mode[mo + 7] * sp;
            double rHy = mode[mo + 8] * cp + mode[mo + 9] * sp;
            g->src[6 * c + 0] = (float)(1 * rHy);
            g->src[6 * c + 1] = (float)(-1 * rHx);
            g->src[6 * c + 2] = 0;
            g->src[6 * c + 3] = (float)((-1 * msign) * rEy);
            g->src[6 * c + 4] = (float)((1 * msign) * rEx);
            g->src[6 * c + 5] = 0;
        }
    }
}

/* Instantaneous overlap -dx^2/2 sum(Hx mEy - Hy mEx + Ey mHx - Ex mHy) on plane cz, with
 * mH -> -conj(mH) when reverse; cum += e^{i reduced phase} * inst (timefield3d.c:3176-3239). */
void orc_mode_overlap(const orc_grid *g, int cx, int cy, int cz, int mi, int mj, double mdisc, double *mode,
                      double phase, int reverse, double cum[2], double inst[2]) {
    orc_mode_power_normalize(mi, mj, mdisc, mode);
    double ratio = g->disc / mdisc;
    double sr = 0.0, si = 0.0;
    for (int i = 0; i < g->I; i++) {
        for (int j = 0; j < g->J; j++) {
            int mx = mode_index_1d(i, cx, ratio, mi);
            int my = mode_index_1d(j, cy, ratio, mj);
            if (mx < 0 || mx >= mi || my < 0 || my >= mj) continue;
            long mo = 2L * 6 * (mx + (long)mi * my);
            double mEx_r = mode[mo + 0], mEx_i = mode[mo + 1];
            double mEy_r = mode[mo + 2], mEy_i = mode[mo + 3];
            double mHx_r = mode[mo + 6], mHx_i = mode[mo + 7];
            double mHy_r = mode[mo + 8], mHy_i = mode[mo + 9];
            if (reverse == 1) {
                mHx_r = 0.0 - mHx_r; mHx_i = 0.0 - (-mHx_i);
                mHy_r = 0.0 - mHy_r; mHy_i = 0.0 - (-mHy_i);
            }
            long c = cidx(g, i, j, cz);
            double Ex = g->f[6 * c + 0], Ey = g->f[6 * c + 1];
            double Hx = g->f[6 * c + 3], Hy = g->f[6 * c + 4];
            double p1r = Hx * mEy_r, p1i = Hx * mEy_i;
            double p2r = 0.0 - Hy * mEx_r, p2i = 0.0 - Hy * mEx_i;
            double p3r = Ey * mHx_r, p3i = Ey * mHx_i;
            double p4r = 0.0 - Ex * mHy_r, p4i = 0.0 - Ex * mHy_i;
            sr = sr + ((p1r + p2r) + (p3r + p4r));
            si = si + ((p1i + p2i) + (p3i + p4i));
        }
    }
    double s = -0.5 * g->disc * g->disc;
    inst[0] = s * sr;
    inst[1] = s * si;
    double tw[2];
    orc_complex_exp(0.0, orc_reduce_phase(phase), tw);
    cum[0] = cum[0] + (tw[0] * inst[0] - tw[1] * inst[1]);
    cum[1] = cum[1] + (tw[0] * inst[1] + tw[1] * inst[0]);
}

/* ------------------------------------------------------------------------------------------ */
/* Geometry and adjoint taper                                                                 */
/* ------------------------------------------------------------------------------------------ */

/* Linear taper rasterised with round(slope*dk) on each edge; ep, mu stored as float in all three
 * components (timefield3d.c:1511-1535). */
void orc_set_taper(const orc_grid *g, double ep, double mu, int start_left, int start_right, int end_left,
                   int end_right, int start_z, int end_z, int bottom, int top) {
    double rs = (double)(end_right - start_right) / ((double)(end_z - start_z));
    double ls = (double)(end_left - start_left) / ((double)(end_z - start_z));
    for (int j = bottom; j < top; j++) {
        for (int k = start_z; k < end_z; k++) {
            int lo = (int)(start_left + round(ls * (k - start_z)));
            double hi = start_right + round(rs * (k - start_z));
            for (int i = lo; i < hi; i++) {
                long c = cidx(g, i, j, k);
                for (int m = 0; m < 3; m++) {
                    g->ep[3 * c + m] = (float)ep;
                    g->mu[3 * c + m] = (float)mu;
                }
            }
        }
    }
}

/* Segment n spans [start + n*L/S, start + (n+1)*L/S) (adjoint3d.c:243-257). */
void orc_taper_insert(const orc_grid *g, double ep, double mu, int seg_num, int start, int end,
                      const int *left, const int *right, int bottom, int top) {
    int L = end - start;
    for (int n = 0; n < seg_num; n++) {
        int s0 = start + (n * L) / seg_num;
        int s1 = start + ((n + 1) * L) / seg_num;
        orc_set_taper(g, ep, mu, left[n], right[n], left[n + 1], right[n + 1], s0, s1, bottom, top);
    }
}

/* Edge pixel of segment n at row k: e[n] + round(int((e[n+1]-e[n])*(k-s0)/(s1-s0)))
 * -- integer division before round (adjoint3d.c:259-297). hist is double[T][L][H][2][6]. */
void orc_taper_store(const orc_grid *g, int seg_num, int start, int end, const int *left, const int *right,
                     int bottom, int top, int timestep, double *hist) {
    int L = end - start, H = top - bottom;
    for (int n = 0; n < seg_num; n++) {
        int s0 = start + (n * L) / seg_num;
        int s1 = start + ((n + 1) * L) / seg_num;
        for (int k = s0; k < s1; k++) {
            for (int j = bottom; j < top; j++) {
                for (int side = 0; side < 2; side++) {
                    const int *e = side == 0 ? right : left;
                    int i = (int)(e[n] + round((double)(((e[n + 1] - e[n]) * (k - s0)) / (s1 - s0))));
                    long c = cidx(g, i, j, k);
                    long h = 6L * (side + 2L * ((j - bottom) + (long)H * ((k - start) + (long)L * timestep)));
                    for (int m = 0; m < 6; m++) hist[h + m] = g->f[6 * c + m];
                }
            }
        }
    }
}

/* g[side,j,k] = sum_t sum_{m<3} fwd*adj (adjoint3d.c:339-373). */
void orc_taper_ep_gradient(int height, int length, int duration, const double *fwd, const double *adj, double *g) {
    for (int side = 0; side < 2; side++) {
        for (int j = 0; j < height; j++) {
            for (int k = 0; k < length; k++) {
                double sum = 0;
                for (int t = 0; t < duration; t++) {
                    long o = 6L * (side + 2L * (j + (long)height * (k + (long)length * t)));
                    for (int m = 0; m < 3; m++) sum += fwd[o + m] * adj[o + m];
                }
                g[side + 2 * (j + height * k)] = sum;
            }
        }
    }
}

/* Hat-weighted projection onto interior control points; left side negated; end points 0
 * (adjoint3d.c:375-418). */
void orc_taper_grad_from_ep(int seg_num, int height, int length, const double *epg, double ep_diff,
                            double *left_grad, double *right_grad) {
    left_grad[0] = 0;
    right_grad[0] = 0;
    for (int n = 1; n < seg_num; n++) {
        left_grad[n] = 0;
        right_grad[n] = 0;
        int sb = ((n - 1) * length) / seg_num;
        int sm = (n * length) / seg_num;
        int sf = ((n + 1) * length) / seg_num;
        for (int k = sb; k < sm; k++) {
            for (int j = 0; j < height; j++) {
                double w = (double)(k - sb) / (double)(sm - sb);
                right_grad[n] += w * epg[0 + 2 * (j + height * k)] * ep_diff;
                left_grad[n] -= w * epg[1 + 2 * (j + height * k)] * ep_diff;
            }
        }
        for (int k = sm; k < sf; k++) {
            for (int j = 0; j < height; j++) {
                double w = (double)(sf - k) / (double)(sf - sm);
                right_grad[n] += w * epg[0 + 2 * (j + height * k)] * ep_diff;
                left_grad[n] -= w * epg[1 + 2 * (j + height * k)] * ep_diff;
            }
        }
    }
    left_grad[seg_num] = 0;
    right_grad[seg_num] = 0;
}

static double asym_norm(int seg_num, const double *lg, const double *rg, double disc, double *dir) {
    double mag = 0;
    for (int i = 0; i < seg_num + 1; i++) {
        mag += rg[i] * rg[i];
        mag += lg[i] * lg[i];
    }
    mag = sqrt(mag) * 0.5 * disc * disc;
    for (int i = 0; i < seg_num + 1; i++) {
        dir[2 * i] = rg[i] * 0.5 * disc * disc / mag;
        dir[2 * i + 1] = lg[i] * 0.5 * disc * disc / mag;
    }
    return mag;
}

/* Independent left/right edge moves with the "zero the smallest component" retry loop and a
 * minimum-width clamp (adjoint3d.c:564-664). */
int orc_taper_update_asymmetric(int seg_num, int *left, int *right, double *left_grad, double *right_grad,
                                double goal, double disc, int min_width) {
    double *dir = malloc(2 * (seg_num + 1) * sizeof(double));
    double mag = asym_norm(seg_num, left_grad, right_grad, disc, dir);
    int unchanged = 1;
    for (int i = 2; i < 2 * seg_num; i++)
        if (round(fabs((goal / mag) * dir[i])) >= 1) unchanged = 0;
    int zeroed = 0;
    while (unchanged == 1 && zeroed < 2 * (seg_num - 2)) {
        double best = 1e10;
        int at = 0;
        for (int i = 2; i < 2 * seg_num; i++) {
            double a = fabs(dir[i]);
            if (a < best && a > 0) { best = a; at = i; }
        }
        if (at % 2 == 0) right_grad[at / 2] = 0;
        else left_grad[at / 2] = 0;
        mag = asym_norm(seg_num, left_grad, right_grad, disc, dir);
        unchanged = 1;
        for (int i = 2; i < 2 * seg_num; i++)
            if (round(fabs((goal / mag) * dir[i])) >= 1) unchanged = 0;
        zeroed++;
    }
    for (int i = 1; i < seg_num; i++) {
        right[i] += (int)(round((goal / mag) * dir[2 * i]));
        left[i] += (int)(round((goal / mag) * dir[2 * i + 1]));
        if (right[i] - left[i] < min_width) {
            right[i] = (right[i] + left[i] + min_width) / 2;
            left[i] = (right[i] + left[i] - min_width) / 2;
        }
    }
    free(dir);
    return unchanged;
}

static double sym_norm(int seg_num, const double *lg, const double *rg, double disc, double *dir) {
    double mag = 0;
    for (int i = 0; i < seg_num + 1; i++) mag += (rg[i] - lg[i]) * (rg[i] - lg[i]);
    mag = sqrt(mag) * 0.5 * disc * disc;
    for (int i = 0; i < seg_num + 1; i++) dir[i] = (rg[i] - lg[i]) * 0.5 * disc * disc / mag;
    return mag;
}

/* Symmetric edge moves (adjoint3d.c:471-562). */
int orc_taper_update_symmetric(int seg_num, int *left, int *right, double *left_grad, double *right_grad,
                               double goal, double disc, int min_width) {
    double *dir = malloc((seg_num + 1) * sizeof(double));
    double mag = sym_norm(seg_num, left_grad, right_grad, disc, dir);
    int unchanged = 1;
    for (int i = 1; i < seg_num; i++)
        if (fabs((goal / mag) * dir[i]) >= 0.5) unchanged = 0;
    int zeroed = 0;
    while (unchanged == 1 && zeroed < seg_num - 1) {
        double best = 1000;
        int at = 0;
        for (int i = 1; i < seg_num; i++) {
            double a = fabs(dir[i]) * 2 * disc;
            if (a < best && a > 0) { best = a; at = i; }
        }
        right_grad[at] = 0;
        left_grad[at] = 0;
        mag = sym_norm(seg_num, left_grad, right_grad, disc, dir);
        unchanged = 1;
        for (int i = 1; i < seg_num; i++)
            if (fabs((goal / mag) * dir[i]) >= 0.5) unchanged = 0;
        zeroed++;
    }
    for (int i = 1; i < seg_num; i++) {
        left[i] -= (int)(round((goal / mag) * dir[i]));
        right[i] += (int)(round((goal / mag) * dir[i]));
        if (right[i] - left[i] < min_width) {
            right[i] = (right[i] + left[i] + min_width) / 2;
            left[i] = (right[i] + left[i] - min_width) / 2;
        }
    }
    free(dir);
    return unchanged;
}

/* ------------------------------------------------------------------------------------------ */
/* Extension: CPML (Roden-Gedney convolutional PML with CFS grading; no reference counterpart) */
/* ------------------------------------------------------------------------------------------ */
/*
 * Units are the reference's normalised ones (c = eps0 = mu0 = 1, lengths in um).  Along an
 * axis of N cells with PML thickness nlo (low) / nhi (high), the depth of a position x is
 *   rho(x) = (nlo - x)/nlo            for x < nlo,
 *   rho(x) = (x - (N-1-nhi))/nhi      for x > N-1-nhi,
 *   0 otherwise (clamped to <= 1);
 * E positions are x = n, H positions x = n + 1/2.  With delta = npml*dx and
 *   sigma_max = -(m+1) ln(R0) / (2 delta),  sigma = sigma_max rho^m,
 *   kappa = 1 + (kappa_max-1) rho^m,  alpha = alpha_max (1 - rho),
 *   b = exp(-(sigma/kappa + alpha) dt),  a = sigma (b-1) / (sigma kappa + kappa^2 alpha).
 * The psi recursion acts on the raw difference d (not divided by dx):
 *   psi <- b psi + a d,  curl term = (d / kappa + psi) / dx.
 * Membership: E position n is in the slab if n < nlo or n >= N-nhi; H position if n < nlo or
 * n >= N-1-nhi (high-side rules only when nhi > 0).  Outside, psi stays 0 and 1/kappa = 1, so
 * the update reduces exactly to the
 * reference interior update.  Boundary cells (index 0 and N-1) are not updated (PEC wall).
 */
void orc_cpml_profile(int N, int nlo, int nhi, const orc_cpml_params *p, double disc, double dt,
                      double *bE, double *aE, double *kE, double *bH, double *aH, double *kH) {
    for (int n = 0; n < N; n++) {
        for (int half = 0; half < 2; half++) {
            double x = n + 0.5 * half;
            double rho = 0.0;
            int npml = 0;
            int in_slab;
            if (half == 0) in_slab = (n < nlo) || (nhi > 0 && n >= N - nhi);
            else in_slab = (n < nlo) || (nhi > 0 && n >= N - 1 - nhi);
            if (in_slab) {
                if (n < nlo) { rho = (nlo - x) / nlo; npml = nlo; }
                else { rho = (x - (N - 1 - nhi)) / nhi; npml = nhi; }
                if (rho > 1.0) rho = 1.0;
                if (rho < 0.0) rho = 0.0;
            }
            double b = 0.0, a = 0.0, kinv = 1.0;
            if (in_slab) {
                double delta = npml * disc;
                double smax = -(p->order + 1) * log(p->R0) / (2.0 * delta);
                double rm = pow(rho, p->order);
                double sigma = smax * rm;
                double kappa = 1.0 + (p->kappa_max - 1.0) * rm;
                double alpha = p->alpha_max * (1.0 - rho);
                b = exp(-(sigma / kappa + alpha) * dt);
                double den = sigma * kappa + kappa * kappa * alpha;
                a = den > 0.0 ? sigma * (b - 1.0) / den : 0.0;
                kinv = 1.0 / kappa;
            }
            if (half == 0) { bE[n] = b; aE[n] = a; kE[n] = kinv; }
            else { bH[n] = b; aH[n] = a; kH[n] = kinv; }
        }
    }
}

static inline int in_slab_E(int n, int N, int nlo, int nhi) { return n < nlo || (nhi > 0 && n >= N - nhi); }
static inline int in_slab_H(int n, int N, int nlo, int nhi) { return n < nlo || (nhi > 0 && n >= N - 1 - nhi); }

void orc_update_E_cpml(const orc_grid *g, const orc_cpml *cp, double dt) {
    const int N[3] = {g->I, g->J, g->K};
    const double inv_dx = 1 / g->disc;
    const long st[3] = {6, 6L * g->I, 6L * g->I * g->J};
#pragma omp parallel for schedule(static)
    for (int k = 1; k < N[2] - 1; k++) {
        for (int j = 1; j < N[1] - 1; j++) {
            for (int i = 1; i < N[0] - 1; i++) {
                const int x[3] = {i, j, k};
                long c = cidx(g, i, j, k);
                double *f = g->f + 6 * c;
                double curl[3];
                for (int a = 0; a < 3; a++) {
                    int b = (a + 1) % 3, cc = (a + 2) % 3;
                    /* first term: d/d_b of H_cc ; second: d/d_cc of H_b (backward differences) */
                    double d1 = f[3 + cc] - f[3 + cc - st[b]];
                    double d2 = f[3 + b] - f[3 + b - st[cc]];
                    double t1 = d1 * cp->kE[b][x[b]];
                    double t2 = d2 * cp->kE[cc][x[cc]];
                    if (in_slab_E(x[b], N[b], cp->nlo[b], cp->nhi[b])) {
                        double *ps = cp->psi[0][a][b] + c;
                        *ps = cp->bE[b][x[b]] * *ps + cp->aE[b][x[b]] * d1;
                        t1 = t1 + *ps;
                    } else {
                        t1 = t1 + 0.0;
                    }
                    if (in_slab_E(x[cc], N[cc], cp->nlo[cc], cp->nhi[cc])) {
                        double *ps = cp->psi[0][a][cc] + c;
                        *ps = cp->bE[cc][x[cc]] * *ps + cp->aE[cc][x[cc]] * d2;
                        t2 = t2 + *ps;
                    } else {
                        t2 = t2 + 0.0;
                    }
                    curl[a] = inv_dx * (t1 - t2);
                }
                for (int m = 0; m < 3; m++) {
                    double ep = g->ep[3 * c + m];
                    double s = g->sig[6 * c + m];
                    double jsrc = g->src[6 * c + m];
                    double inv_den = 1.0 / (ep / dt + 0.5 * s);
                    double keep = ep / dt - 0.5 * s;
                    f[m] = inv_den * (keep * f[m] + curl[m] - jsrc);
                }
            }
        }
    }
}

void orc_update_H_cpml(const orc_grid *g, const orc_cpml *cp, double dt) {
    const int N[3] = {g->I, g->J, g->K};
    const double inv_dx = 1 / g->disc;
    const long st[3] = {6, 6L * g->I, 6L * g->I * g->J};
#pragma omp parallel for schedule(static)
    for (int k = 1; k < N[2] - 1; k++) {
        for (int j = 1; j < N[1] - 1; j++) {
            for (int i = 1; i < N[0] - 1; i++) {
                const int x[3] = {i, j, k};
                long c = cidx(g, i, j, k);
                double *f = g->f + 6 * c;
                double curl[3];
                for (int a = 0; a < 3; a++) {
                    int b = (a + 1) % 3, cc = (a + 2) % 3;
                    /* forward differences: d/d_b of E_cc minus d/d_cc of E_b */
                    double d1 = f[cc + st[b]] - f[cc];
                    double d2 = f[b + st[cc]] - f[b];
                    double t1 = d1 * cp->kH[b][x[b]];
                    double t2 = d2 * cp->kH[cc][x[cc]];
                    if (in_slab_H(x[b], N[b], cp->nlo[b], cp->nhi[b])) {
                        double *ps = cp->psi[1][a][b] + c;
                        *ps = cp->bH[b][x[b]] * *ps + cp->aH[b][x[b]] * d1;
                        t1 = t1 + *ps;
                    } else {
                        t1 = t1 + 0.0;
                    }
                    if (in_slab_H(x[cc], N[cc], cp->nlo[cc], cp->nhi[cc])) {
                        double *ps = cp->psi[1][a][cc] + c;
                        *ps = cp->bH[cc][x[cc]] * *ps + cp->aH[cc][x[cc]] * d2;
                        t2 = t2 + *ps;
                    } else {
                        t2 = t2 + 0.0;
                    }
                    curl[a] = inv_dx * (t1 - t2);
                }
                for (int m = 0; m < 3; m++) {
                    double h = f[3 + m];
                    double msrc = g->src[6 * c + 3 + m] + h * g->sig[6 * c + 3 + m];
                    double mu = g->mu[3 * c + m];
                    double s = g->sig[6 * c + 3 + m];
                    double inv_den = 1.0 / (mu / dt + 0.5 * s);
                    double keep = mu / dt - 0.5 * s;
                    f[3 + m] = inv_den * (keep * h - curl[m] - msrc);
                }
            }
        }
    }
}

void orc_step_cpml(const orc_grid *g, const orc_cpml *c, double dt) {
    orc_update_E_cpml(g, c, dt);
    orc_update_H_cpml(g, c, dt);
}

/* ------------------------------------------------------------------------------------------ */
/* Extension: multi-frequency running DFT of a scalar                                          */
/* ------------------------------------------------------------------------------------------ */

/* acc[f] += value * e^{i omega_f t} with libm cos/sin of the unreduced phase omega_f*t. */
void orc_dft_scalar(int nfreq, const double *omega, double t, double value, double *acc) {
    for (int q = 0; q < nfreq; q++) {
        double ph = omega[q] * t;
        acc[2 * q] += value * cos(ph);
        acc[2 * q + 1] += value * sin(ph);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* extension: DFT boxes and the DFT-domain permittivity gradient                              */
/* ------------------------------------------------------------------------------------------ */

void orc_dft_box(const orc_grid *g, const int box[6], int mask, int nfreq, const double *omega, double t,
                 double *acc) {
    const long ni = box[1] - box[0], nj = box[3] - box[2], nk = box[5] - box[4], nb = ni * nj * nk;
    for (int f = 0; f < nfreq; f++) {
        const double tr = cos(omega[f] * t), ti = sin(omega[f] * t);
        int slot = 0;
        for (int m = 0; m < 6; m++) {
            if (!(mask >> m & 1)) continue;
            double *a = acc + 2 * ((long)slot * nfreq + f) * nb;
            for (long q = 0; q < nb; q++) {
                const int i = box[0] + (int)(q % ni), j = box[2] + (int)((q / ni) % nj), k = box[4] + (int)(q / (ni * nj));
                const double v = g->f[6 * cidx(g, i, j, k) + m];
                a[2 * q] += v * tr;
                a[2 * q + 1] += v * ti;
            }
            slot++;
        }
    }
}

/* The overlap inst = -dx^2/2 sum(Hx mEy - Hy mEx + Ey mHx' - Ex mHy') is w_E.E + w_H.H with
 * w_E = dx^2/2 (mHy', -mHx', 0), w_H = -dx^2/2 (mEy, -mEx, 0).  In the frequency domain of the
 * semi-implicit scheme (z = e^{i omega dt}, u = 1/z - 1) the monitor reads g.E with
 * g = w_E - (dt/(z u)) C_H mu^-1 w_H, and a magnetic source M acts on E like the electric source
 * (dt/u) C_H mu^-1 M.  The adjoint source J = w_E, M = -w_H / z therefore drives g, the field it
 * radiates equals (by the symmetry of the discrete operator) the response of the monitor to a unit
 * current at every cell, and the derivative of the monitor with respect to eps follows from it
 * (see orc_dft_gradient). */
void orc_overlap_adjoint_source(const orc_grid *g, int cx, int cy, int cz, int mi, int mj, double mdisc,
                                const double *mode, int reverse, double omega_dt, double amplitude, double phase) {
    const double ratio = g->disc / mdisc, h = 0.5 * g->disc * g->disc;
    const double zr = cos(omega_dt), zi = -sin(omega_dt);
    const double cp = cos(phase), sp = sin(phase);
    for (int i = 0; i < g->I; i++) {
        for (int j = 0; j < g->J; j++) {
            int mx = mode_index_1d(i, cx, ratio, mi);
            int my = mode_index_1d(j, cy, ratio, mj);
            if (mx < 0 || mx >= mi || my < 0 || my >= mj) continue;
            const long c = cidx(g, i, j, cz);
            const double *md = mode + 12L * (mx + (long)mi * my);
            double hxr = md[6], hxi = md[7], hyr = md[8], hyi = md[9];
            if (reverse == 1) {
                hxr = -hxr;
                hyr = -hyr;
            }
            const double J[3][2] = {{h * hyr, h * hyi}, {-h * hxr, -h * hxi}, {0, 0}};
            const double Mr[3][2] = {{h * md[2], h * md[3]}, {-h * md[0], -h * md[1]}, {0, 0}};
            for (int m = 0; m < 3; m++) {
                const double Mre = zr * Mr[m][0] - zi * Mr[m][1], Mim = zr * Mr[m][1] + zi * Mr[m][0];
                g->src[6 * c + m] = (float)(amplitude * (J[m][0] * cp + J[m][1] * sp));
                g->src[6 * c + 3 + m] = (float)(amplitude * (Mre * cp + Mim * sp));
            }
        }
    }
}

void orc_dft_gradient(long nb, int fmask, int fnf, const double *F, int amask, int anf, const double *A, int f,
                      double wr, double wi, double *grad) {
    for (long q = 0; q < nb; q++) {
        double sr = 0, si = 0;
        int fs = 0, as = 0;
        for (int m = 0; m < 6; m++) {
            const int inf = fmask >> m & 1, ina = amask >> m & 1;
            if (m < 3 && inf && ina) {
                const double *x = F + 2 * (((long)fs * fnf + f) * nb + q), *y = A + 2 * (((long)as * anf + f) * nb + q);
                sr += x[0] * y[0] - x[1] * y[1];
                si += x[0] * y[1] + x[1] * y[0];
            }
            fs += inf;
            as += ina;
        }
        grad[q] = wr * sr - wi * si;
    }
}
