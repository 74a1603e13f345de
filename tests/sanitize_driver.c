/* Driver for the sanitizer build (tests/test_sanitizers.py): exercises the pure host parts of the
 * library (complex helpers, small eigen-solver, mode solver, mode file round trip) and the FP64
 * oracle's Yee and CPML steps under AddressSanitizer / UndefinedBehaviorSanitizer.  Exits 0 when
 * every result is finite and the round trip is exact; any sanitizer report aborts first. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "complex.h"
#include "fdtd_oracle.h"
#include "matrix.h"
#include "mode.h"

static unsigned long long lcg = 12345;
static double urand(void) {
    lcg = lcg * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(lcg >> 11) / 9007199254740992.0 * 2.0 - 1.0;
}

static int eigen_part(void) {
    const int n = 12;
    matrix A = matrix_new(n, n);
    for (int q = 0; q < n * n; q++) A.m[q] = complex_from_doubles(urand(), urand());
    matrix D = matrix_diagonalize(A, 1e-3, 1000);
    matrix x = matrix_eigenvector(A, D.m[0]);
    int ok = isfinite(D.m[0].re) && isfinite(x.m[0].re);
    matrix_free(x);
    matrix_free(D);
    matrix_free(A);
    return ok;
}

static int mode_part(void) {
    const int mi = 26, mj = 20;
    mode md = mode_new(mi, mj, 0.04);
    for (int y = 0; y < mj; y++)
        for (int x = 0; x < mi; x++)
            for (int k = 0; k < 3; k++) {
                double e = (y >= 8 && y < 12 && x >= 7 && x < 19) ? 3.48 * 3.48 : 1.44 * 1.44;
                md.ep.m[3 * (x + mi * y) + k] = complex_from_doubles(e, 0);
                md.mu.m[3 * (x + mi * y) + k] = complex_from_doubles(1, 0);
            }
    mode_set_pml(md, 0.5, 2, 3);
    complex b2 = mode_solve_beta_squared(&md, 2 * 3.14159265358979 / 1.55, 1e-8, NULL, 0, 20, 2, 0);
    int ok = isfinite(b2.re) && b2.re > 0;
    mode_power_normalize(md);
    FILE *f = tmpfile();
    if (!f) return 0;
    mode_file_print(f, md, 2);
    rewind(f);
    mode back = mode_file_load(f);
    fclose(f);
    ok = ok && back.i == mi && back.j == mj;
    for (int q = 0; q < 6 * mi * mj && ok; q++) ok = fabs(back.fields.m[q].re - md.fields.m[q].re) < 1e-5;
    mode_free(back);
    mode_free(md);
    return ok;
}

static int oracle_part(void) {
    const int I = 14, J = 11, K = 13;
    const long n = (long)I * J * K;
    orc_grid g = {I, J, K, 0.02, malloc(sizeof(float) * 3 * n), malloc(sizeof(float) * 3 * n),
                  calloc(6 * n, sizeof(float)), malloc(sizeof(double) * 6 * n), calloc(6 * n, sizeof(double)),
                  calloc(6 * n, sizeof(float))};
    for (long q = 0; q < 3 * n; q++) g.ep[q] = (float)(2.0 + urand()), g.mu[q] = 1.0f;
    for (long q = 0; q < 6 * n; q++) g.f[q] = urand();
    const double dt = 0.99 * 0.02 / sqrt(3.0);
    for (int s = 0; s < 3; s++) orc_step(&g, dt);
    orc_cpml_params p = {{3, 2, 0, 3, 2, 4}, 3, 1e-6, 1.5, 0.1};
    orc_cpml c;
    memset(&c, 0, sizeof c);
    const int N[3] = {I, J, K};
    double *tabs[3][6];
    for (int a = 0; a < 3; a++) {
        c.nlo[a] = p.npml[2 * a];
        c.nhi[a] = p.npml[2 * a + 1];
        for (int t = 0; t < 6; t++) tabs[a][t] = calloc(N[a], sizeof(double));
        orc_cpml_profile(N[a], c.nlo[a], c.nhi[a], &p, g.disc, dt, tabs[a][0], tabs[a][1], tabs[a][2], tabs[a][3],
                         tabs[a][4], tabs[a][5]);
        c.bE[a] = tabs[a][0], c.aE[a] = tabs[a][1], c.kE[a] = tabs[a][2];
        c.bH[a] = tabs[a][3], c.aH[a] = tabs[a][4], c.kH[a] = tabs[a][5];
    }
    for (int f = 0; f < 2; f++)
        for (int m = 0; m < 3; m++)
            for (int a = 0; a < 3; a++) c.psi[f][m][a] = calloc(n, sizeof(double));
    for (int s = 0; s < 3; s++) orc_step_cpml(&g, &c, dt);
    int ok = 1;
    for (long q = 0; q < 6 * n && ok; q++) ok = isfinite(g.f[q]);
    for (int f = 0; f < 2; f++)
        for (int m = 0; m < 3; m++)
            for (int a = 0; a < 3; a++) free(c.psi[f][m][a]);
    for (int a = 0; a < 3; a++)
        for (int t = 0; t < 6; t++) free(tabs[a][t]);
    free(g.ep), free(g.mu), free(g.sig), free(g.f), free(g.old), free(g.src);
    return ok;
}

int main(void) {
    int e = eigen_part(), m = mode_part(), o = oracle_part();
    printf("eigen %d mode %d oracle %d\n", e, m, o);
    return e && m && o ? 0 : 1;
}
