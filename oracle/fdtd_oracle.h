/*
 * fdtd_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A clean-room FP64 CPU restatement of the reference FDTD hot path, used by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py as the CHECKER.  No product
 * code links, loads or calls it.  Every routine cites the reference file:line whose behaviour
 * it restates; the reference-compatible routines are pinned bit-for-bit against the compiled
 * reference (oracle/_ref) through the committed fixtures in tests/golden/.
 *
 * Layout (same as the reference, timefield3d.h:15-27): cell c = i + I*(j + J*k), x fastest;
 * fields/old are double[6n] ordered Ex,Ey,Ez,Hx,Hy,Hz per cell; ep/mu are float[3n];
 * sig/src are float[6n] (electric at m=0..2, magnetic at m=3..5).  Complex arrays are
 * interleaved (re, im) doubles.
 *
 * Extensions that the reference does not have (CPML, multi-frequency DFT, DFT-domain
 * permittivity gradient) follow the formulas written next to each routine; they are pinned
 * by invariants in tests/ (see the header of fdtd_oracle.c for the list).
 */
#ifndef FDTD_ORACLE_H
#define FDTD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int I, J, K;
    double disc;
    float *ep, *mu, *sig;
    double *f, *old;
    float *src;
} orc_grid;

/* ---- complex helpers (complex.c semantics) ---- */
void orc_complex_exp(double re, double im, double out[2]);
double orc_complex_angle(double re, double im);
void orc_complex_sqrt(double re, double im, double out[2]);
double orc_reduce_phase(double phase);

/* ---- Yee step, reference mode (Mur) ---- */
void orc_snapshot_shells(const orc_grid *g);
void orc_update_E(const orc_grid *g, double dt);
void orc_update_H(const orc_grid *g, double dt);
void orc_mur_E(const orc_grid *g, double dt);
void orc_mur_H(const orc_grid *g, double dt);
void orc_step(const orc_grid *g, double dt);

/* ---- running DFT (single frequency, reference semantics) ---- */
void orc_dft_volume(const orc_grid *g, double phase, double *F);
void orc_dft_y_slice(const orc_grid *g, double phase, int j, double *F);

/* ---- modes ---- */
void orc_mode_power_inner(int mi, int mj, double mdisc, const double *psi, const double *phi, double out[2]);
void orc_mode_power_normalize(int mi, int mj, double mdisc, double *mode);
void orc_mode_source(const orc_grid *g, int cx, int cy, int cz, int mi, int mj, double mdisc, double *mode,
                     double phase, double amplitude, int reverse);
void orc_mode_source_adjoint(const orc_grid *g, int cx, int cy, int cz, int mi, int mj, double mdisc, double *mode,
                             double phase, int reverse);
void orc_mode_overlap(const orc_grid *g, int cx, int cy, int cz, int mi, int mj, double mdisc, double *mode,
                      double phase, int reverse, double cum[2], double inst[2]);

/* ---- geometry (host setters) ---- */
void orc_set_taper(const orc_grid *g, double ep, double mu, int start_left, int start_right, int end_left,
                   int end_right, int start_z, int end_z, int bottom, int top);

/* ---- adjoint taper (adjoint3d.c semantics) ---- */
void orc_taper_insert(const orc_grid *g, double ep, double mu, int seg_num, int start, int end,
                      const int *left, const int *right, int bottom, int top);
void orc_taper_store(const orc_grid *g, int seg_num, int start, int end, const int *left, const int *right,
                     int bottom, int top, int timestep, double *hist);
void orc_taper_ep_gradient(int height, int length, int duration, const double *fwd, const double *adj, double *g);
void orc_taper_grad_from_ep(int seg_num, int height, int length, const double *epg, double ep_diff,
                            double *left_grad, double *right_grad);
int orc_taper_update_asymmetric(int seg_num, int *left, int *right, double *left_grad, double *right_grad,
                                double improvement_goal, double disc, int min_width);
int orc_taper_update_symmetric(int seg_num, int *left, int *right, double *left_grad, double *right_grad,
                               double improvement_goal, double disc, int min_width);

/* ---- extension: CPML (no reference counterpart) ---- */
typedef struct {
    int npml[6];        /* cells, faces x-lo, x-hi, y-lo, y-hi, z-lo, z-hi (0 = no PML on that face) */
    int order;          /* polynomial grading order m */
    double R0;          /* target normal-incidence reflection */
    double kappa_max;
    double alpha_max;
} orc_cpml_params;

typedef struct {
    int nlo[3], nhi[3];  /* slab thickness per axis (cells) */
    /* per-axis 1-D profile tables, length N of that axis, for E positions (n) and H positions (n+1/2) */
    double *bE[3], *aE[3], *kE[3];
    double *bH[3], *aH[3], *kH[3];
    /* full-volume psi arrays (zero outside the slabs), 12 of them: index [field(0=E,1=H)][comp][deriv axis] */
    double *psi[2][3][3];
} orc_cpml;

void orc_cpml_profile(int N, int nlo, int nhi, const orc_cpml_params *p, double disc, double dt,
                      double *bE, double *aE, double *kE, double *bH, double *aH, double *kH);
void orc_update_E_cpml(const orc_grid *g, const orc_cpml *c, double dt);
void orc_update_H_cpml(const orc_grid *g, const orc_cpml *c, double dt);
void orc_step_cpml(const orc_grid *g, const orc_cpml *c, double dt);

/* ---- extension: multi-frequency running DFT of a scalar and of a plane ---- */
void orc_dft_scalar(int nfreq, const double *omega, double t, double value, double *acc);

/* ---- extension: DFT boxes and the DFT-domain permittivity gradient ---- */
/* acc[slot][f][q] += f_m(cell q) e^{i omega_f t} for each component m in mask (slot = rank of m),
 * box [b0,b1) x [b2,b3) x [b4,b5), q x fastest; acc interleaved complex. */
void orc_dft_box(const orc_grid *g, const int box[6], int mask, int nfreq, const double *omega, double t,
                 double *acc);
/* Exact adjoint source of an overlap monitor (mode already power-normalised): writes plane cz of
 * the sources with Re[(J, M) amplitude e^{-i phase}], J = dx^2/2 (mHy', -mHx', 0),
 * M = e^{-i omega_dt} dx^2/2 (mEy, -mEx, 0), mH' = -conj(mH) when reverse. */
void orc_overlap_adjoint_source(const orc_grid *g, int cx, int cy, int cz, int mi, int mj, double mdisc,
                                const double *mode, int reverse, double omega_dt, double amplitude, double phase);
/* grad[q] = Re[(wr + i wi) sum_{m<3 in both masks} F_m(q) A_m(q)] at frequency index f. */
void orc_dft_gradient(long nb, int fmask, int fnf, const double *F, int amask, int anf, const double *A, int f,
                      double wr, double wi, double *grad);

#ifdef __cplusplus
}
#endif
#endif
