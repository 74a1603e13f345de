/* Dense complex matrices of the drop-in surface (matrix.h): row-major, entry (r, c) at
 * m[c + r*j] (matrix.c:4-60).  Dimension mismatches print a message and abort, as in the
 * reference (matrix.c:443-447). */
#include "matrix.h"

#include <string.h>

static void need(int ok, const char *what) {
    if (!ok) {
        printf("%s\n", what);
        abort();
    }
}

matrix matrix_new(int i, int j) {
    matrix A;
    A.i = i;
    A.j = j;
    A.m = calloc((size_t)i * j > 0 ? (size_t)i * j : 1, sizeof(complex));
    need(A.m != NULL, "matrix_new: out of memory");
    return A;
}

void matrix_free(matrix A) { free(A.m); }

void matrix_print(matrix A) {
    for (int r = 0; r < A.i; r++) {
        for (int c = 0; c < A.j; c++) printf("%f + i%f\t\t", A.m[c + r * A.j].re, A.m[c + r * A.j].im);
        printf("\n");
    }
}

matrix matrix_copy(matrix A) {
    matrix B = matrix_new(A.i, A.j);
    memcpy(B.m, A.m, (size_t)A.i * A.j * sizeof(complex));
    return B;
}

void matrix_copy_in_place(matrix A, matrix B) {
    need(A.i == B.i && A.j == B.j, "matrix dimensions do not match for copy!");
    memcpy(A.m, B.m, (size_t)A.i * A.j * sizeof(complex));
}

matrix matrix_get_column(matrix A, int j) {
    matrix c = matrix_new(A.i, 1);
    for (int r = 0; r < A.i; r++) c.m[r] = A.m[j + r * A.j];
    return c;
}

matrix matrix_get_row(matrix A, int i) {
    matrix r = matrix_new(1, A.j);
    for (int c = 0; c < A.j; c++) r.m[c] = A.m[c + i * A.j];
    return r;
}

void matrix_set_column(matrix A, matrix col, int j) {
    need(col.i == A.i && col.j == 1, "column has wrong dimensions!");
    for (int r = 0; r < A.i; r++) A.m[j + r * A.j] = col.m[r];
}

void matrix_set_row(matrix A, matrix row, int i) {
    need(row.j == A.j && row.i == 1, "row has wrong dimensions!");
    for (int c = 0; c < A.j; c++) A.m[c + i * A.j] = row.m[c];
}

/* A without row i and column j (matrix.c:162).  The reference's bounds check compares the row
 * index with both dimensions and never looks at the column index; here each index is checked
 * against its own dimension, with the reference's message. */
matrix matrix_get_submatrix(matrix A, int i, int j) {
    need(i >= 0 && j >= 0 && i < A.i && j < A.j, "Row or column is out of bounds for getting minor matrix!");
    matrix M = matrix_new(A.i - 1, A.j - 1);
    for (int r = 0, rr = 0; r < A.i; r++) {
        if (r == i) continue;
        for (int c = 0, cc = 0; c < A.j; c++) {
            if (c == j) continue;
            M.m[cc + rr * M.j] = A.m[c + r * A.j];
            cc++;
        }
        rr++;
    }
    return M;
}

/* Laplace expansion along the first row, sum_c (-1)^c A[0][c] det(minor(0, c)), accumulated in
 * column order: the reference's definition (matrix.c:200), used up to 4 x 4 where it costs at
 * most 40 products and reproduces the reference's rounding exactly. */
static complex det_laplace(const complex *a, int n) {
    if (n == 1) return a[0];
    complex minor[3 * 3];
    complex det = complex_zero();
    double sign = 1.0;
    for (int c = 0; c < n; c++) {
        for (int r = 1; r < n; r++)
            for (int q = 0, qq = 0; q < n; q++)
                if (q != c) minor[qq++ + (r - 1) * (n - 1)] = a[q + r * n];
        complex t = complex_mul(det_laplace(minor, n - 1), a[c]);
        det = complex_add(det, complex_smul(sign, t));
        sign = -sign;
    }
    return det;
}

/* Determinant of a square matrix (matrix.c:200).  Above 4 x 4 the reference's recursion costs
 * n! products; here LU factorisation with partial pivoting (Golub & Van Loan 3.4): the product
 * of the pivots times the sign of the row permutation, n^3/3 complex multiply-adds.  An exactly
 * zero pivot column gives 0. */
complex matrix_determinant(matrix A) {
    need(A.i == A.j, "A is not square for finding determinant!");
    const int n = A.i;
    if (n <= 4) return det_laplace(A.m, n);
    complex *U = malloc((size_t)n * n * sizeof(complex));
    need(U != NULL, "out of memory");
    memcpy(U, A.m, (size_t)n * n * sizeof(complex));
    complex det = complex_from_doubles(1, 0);
    for (int col = 0; col < n; col++) {
        int piv = col;
        for (int r = col + 1; r < n; r++)
            if (complex_mag(U[col + r * n]) > complex_mag(U[col + piv * n])) piv = r;
        if (complex_mag(U[col + piv * n]) == 0.0) {
            free(U);
            return complex_zero();
        }
        if (piv != col) {
            for (int c = col; c < n; c++) {
                const complex t = U[c + col * n];
                U[c + col * n] = U[c + piv * n];
                U[c + piv * n] = t;
            }
            det = complex_smul(-1.0, det);
        }
        const complex inv = complex_reciprocal(U[col + col * n]);
        for (int r = col + 1; r < n; r++) {
            const complex f = complex_mul(U[col + r * n], inv);
            for (int c = col + 1; c < n; c++) U[c + r * n] = complex_sub(U[c + r * n], complex_mul(f, U[c + col * n]));
        }
        det = complex_mul(det, U[col + col * n]);
    }
    free(U);
    return det;
}

complex matrix_trace(matrix A) {
    need(A.i == A.j, "trace needs a square matrix!");
    complex t = complex_zero();
    for (int r = 0; r < A.i; r++) t = complex_add(t, A.m[r + r * A.j]);
    return t;
}

matrix matrix_hermitian(matrix A) {
    matrix H = matrix_new(A.j, A.i);
    for (int r = 0; r < A.i; r++)
        for (int c = 0; c < A.j; c++) H.m[r + c * A.i] = complex_conj(A.m[c + r * A.j]);
    return H;
}

matrix matrix_smul(double s, matrix A) {
    matrix B = matrix_new(A.i, A.j);
    for (int q = 0; q < A.i * A.j; q++) B.m[q] = complex_smul(s, A.m[q]);
    return B;
}

matrix matrix_complex_smul(complex s, matrix A) {
    matrix B = matrix_new(A.i, A.j);
    for (int q = 0; q < A.i * A.j; q++) B.m[q] = complex_mul(s, A.m[q]);
    return B;
}

matrix matrix_add(matrix A, matrix B) {
    need(A.i == B.i && A.j == B.j, "matrix dimensions do not match for addition!");
    matrix C = matrix_new(A.i, A.j);
    for (int q = 0; q < A.i * A.j; q++) C.m[q] = complex_add(A.m[q], B.m[q]);
    return C;
}

matrix matrix_sub(matrix A, matrix B) {
    need(A.i == B.i && A.j == B.j, "matrix dimensions do not match for subtraction!");
    matrix C = matrix_new(A.i, A.j);
    for (int q = 0; q < A.i * A.j; q++) C.m[q] = complex_sub(A.m[q], B.m[q]);
    return C;
}

void matrix_smul_in_place(double s, matrix A) {
    for (int q = 0; q < A.i * A.j; q++) A.m[q] = complex_smul(s, A.m[q]);
}

void matrix_complex_smul_in_place(complex s, matrix A) {
    for (int q = 0; q < A.i * A.j; q++) A.m[q] = complex_mul(s, A.m[q]);
}

void matrix_add_in_place(matrix A, matrix B) {
    need(A.i == B.i && A.j == B.j, "matrix dimensions do not match for addition!");
    for (int q = 0; q < A.i * A.j; q++) A.m[q] = complex_add(A.m[q], B.m[q]);
}

void matrix_sub_in_place(matrix A, matrix B) {
    need(A.i == B.i && A.j == B.j, "matrix dimensions do not match for subtraction!");
    for (int q = 0; q < A.i * A.j; q++) A.m[q] = complex_sub(A.m[q], B.m[q]);
}

matrix matrix_mul(matrix A, matrix B) {
    need(A.j == B.i, "matrix inner dimensions do not match for multiplication!");
    matrix C = matrix_new(A.i, B.j);
    for (int r = 0; r < A.i; r++)
        for (int c = 0; c < B.j; c++) {
            complex s = complex_zero();
            for (int k = 0; k < A.j; k++) s = complex_add(s, complex_mul(A.m[k + r * A.j], B.m[c + k * B.j]));
            C.m[c + r * B.j] = s;
        }
    return C;
}

/* Gaussian elimination with partial pivoting on [A | B]; returns X with A X = B. */
matrix matrix_linear_solve(matrix A, matrix B) {
    need(A.i == A.j && B.i == A.i, "matrix dimensions do not match for linear solve!");
    int n = A.i, nr = B.j;
    matrix M = matrix_copy(A), X = matrix_copy(B);
    for (int col = 0; col < n; col++) {
        int piv = col;
        double best = complex_mag(M.m[col + col * n]);
        for (int r = col + 1; r < n; r++) {
            double v = complex_mag(M.m[col + r * n]);
            if (v > best) { best = v; piv = r; }
        }
        need(best > 0, "matrix is singular!");
        if (piv != col) {
            for (int c = 0; c < n; c++) {
                complex t = M.m[c + col * n]; M.m[c + col * n] = M.m[c + piv * n]; M.m[c + piv * n] = t;
            }
            for (int c = 0; c < nr; c++) {
                complex t = X.m[c + col * nr]; X.m[c + col * nr] = X.m[c + piv * nr]; X.m[c + piv * nr] = t;
            }
        }
        complex inv = complex_reciprocal(M.m[col + col * n]);
        for (int r = col + 1; r < n; r++) {
            complex f = complex_mul(M.m[col + r * n], inv);
            for (int c = col; c < n; c++)
                M.m[c + r * n] = complex_sub(M.m[c + r * n], complex_mul(f, M.m[c + col * n]));
            for (int c = 0; c < nr; c++)
                X.m[c + r * nr] = complex_sub(X.m[c + r * nr], complex_mul(f, X.m[c + col * nr]));
        }
    }
    for (int r = n - 1; r >= 0; r--) {
        complex inv = complex_reciprocal(M.m[r + r * n]);
        for (int c = 0; c < nr; c++) {
            complex s = X.m[c + r * nr];
            for (int k = r + 1; k < n; k++) s = complex_sub(s, complex_mul(M.m[k + r * n], X.m[c + k * nr]));
            X.m[c + r * nr] = complex_mul(s, inv);
        }
    }
    matrix_free(M);
    return X;
}

matrix matrix_inverse(matrix A) {
    matrix Id = matrix_new(A.i, A.i);
    for (int r = 0; r < A.i; r++) Id.m[r + r * A.i] = complex_from_doubles(1, 0);
    matrix X = matrix_linear_solve(A, Id);
    matrix_free(Id);
    return X;
}

/* Modified Gram-Schmidt on the columns, in place. */
void matrix_gram_schmidt(matrix A) {
    for (int j = 0; j < A.j; j++) {
        matrix cur = matrix_get_column(A, j);
        for (int p = 0; p < j; p++) {
            matrix prev = matrix_get_column(A, p);
            complex ip = matrix_vector_inner_product(prev, cur);
            for (int r = 0; r < A.i; r++) cur.m[r] = complex_sub(cur.m[r], complex_mul(ip, prev.m[r]));
            matrix_free(prev);
        }
        double nrm = matrix_vector_norm(cur);
        if (nrm > 0) matrix_smul_in_place(1.0 / nrm, cur);
        matrix_set_column(A, cur, j);
        matrix_free(cur);
    }
}

complex matrix_vector_inner_product(matrix A, matrix B) {
    need(A.j == 1 && B.j == 1 && A.i == B.i, "inner product needs column vectors of equal length!");
    complex s = complex_zero();
    for (int r = 0; r < A.i; r++) s = complex_add(s, complex_mul(complex_conj(A.m[r]), B.m[r]));
    return s;
}

double matrix_vector_norm(matrix A) { return sqrt(complex_mag(matrix_vector_inner_product(A, A))); }

matrix matrix_vector_normalize(matrix A) {
    matrix B = matrix_copy(A);
    matrix_vector_normalize_in_place(B);
    return B;
}

void matrix_vector_normalize_in_place(matrix A) {
    double n = matrix_vector_norm(A);
    matrix_smul_in_place(1.0 / n, A);
}

/* ---- eigenvalues of a small dense complex matrix (matrix.c:672-871, 1114-1158) ----
 * The reference reaches a triangular form with an unshifted QL iteration and stops when every
 * entry above the diagonal is below tol x the largest diagonal entry.  Here: Householder
 * reduction to upper Hessenberg form, then single-shift QR steps with Wilkinson shifts and
 * deflation of negligible sub-diagonal entries (Golub & Van Loan, Matrix Computations, 7.4-7.5),
 * so the eigenvalues are converged to working precision; tol is kept in the signature and
 * max_iterations bounds the QR steps per eigenvalue.  Returns the diagonal matrix of the
 * eigenvalues (upper-left to lower-right in deflation order, last ones first). */

static complex cplx(double re, double im) { return complex_from_doubles(re, im); }
static double cabs1(complex z) { return fabs(z.re) + fabs(z.im); }

/* Givens rotation G = [[c, s], [-conj(s), c]] (c real) with G (a, b)^T = (r, 0)^T */
static void givens(complex a, complex b, double *c, complex *s) {
    const double na = complex_mag(a), nb = complex_mag(b);
    if (nb == 0.0) {
        *c = 1.0;
        *s = complex_zero();
        return;
    }
    if (na == 0.0) {
        *c = 0.0;
        *s = complex_conj(complex_smul(1.0 / nb, b));
        return;
    }
    const double r = hypot(na, nb);
    *c = na / r;
    *s = complex_smul(1.0 / r, complex_mul(complex_smul(1.0 / na, a), complex_conj(b)));
}

static void hessenberg(complex *H, int n) {
    for (int k = 0; k < n - 2; k++) {
        double alpha = 0.0;
        for (int r = k + 1; r < n; r++) alpha += complex_mag(H[k + r * n]) * complex_mag(H[k + r * n]);
        alpha = sqrt(alpha);
        if (alpha == 0.0) continue;
        /* v = x + e^{i arg x0} |x| e0, reflector I - 2 v v^H / (v^H v) */
        const complex x0 = H[k + (k + 1) * n];
        const double m0 = complex_mag(x0);
        const complex ph = m0 > 0 ? complex_smul(1.0 / m0, x0) : cplx(1, 0);
        complex *v = calloc((size_t)n, sizeof(complex));
        for (int r = k + 1; r < n; r++) v[r] = H[k + r * n];
        v[k + 1] = complex_add(v[k + 1], complex_smul(alpha, ph));
        double vv = 0.0;
        for (int r = k + 1; r < n; r++) vv += complex_mag(v[r]) * complex_mag(v[r]);
        if (vv == 0.0) {
            free(v);
            continue;
        }
        for (int c = 0; c < n; c++) {  /* H <- (I - 2 v v^H / vv) H */
            complex d = complex_zero();
            for (int r = k + 1; r < n; r++) d = complex_add(d, complex_mul(complex_conj(v[r]), H[c + r * n]));
            d = complex_smul(2.0 / vv, d);
            for (int r = k + 1; r < n; r++) H[c + r * n] = complex_sub(H[c + r * n], complex_mul(v[r], d));
        }
        for (int r = 0; r < n; r++) {  /* H <- H (I - 2 v v^H / vv) */
            complex d = complex_zero();
            for (int c = k + 1; c < n; c++) d = complex_add(d, complex_mul(H[c + r * n], v[c]));
            d = complex_smul(2.0 / vv, d);
            for (int c = k + 1; c < n; c++) H[c + r * n] = complex_sub(H[c + r * n], complex_mul(d, complex_conj(v[c])));
        }
        for (int r = k + 2; r < n; r++) H[k + r * n] = complex_zero();
        free(v);
    }
}

/* Eigenvalues of the n x n matrix a into ev[0..n) (deflation order); 0 when some eigenvalue took
 * more than max(max_iterations, 30) QR steps (the ones not deflated are then the diagonal entries
 * reached so far). */
static int schur_eigenvalues(const complex *a, int n, int max_iterations, complex *ev) {
    complex *H = malloc((size_t)n * n * sizeof(complex));
    need(H != NULL, "out of memory");
    memcpy(H, a, (size_t)n * n * sizeof(complex));
    hessenberg(H, n);
    int converged = 1;
    const double eps = 2.220446049250313e-16;
    int hi = n - 1, its = 0;
    const int cap = max_iterations > 30 ? max_iterations : 30;
    while (hi > 0) {
        int lo = hi;
        while (lo > 0) {  /* find the top of the unreduced block ending at hi */
            const double scale = cabs1(H[lo + lo * n]) + cabs1(H[(lo - 1) + (lo - 1) * n]);
            if (cabs1(H[(lo - 1) + lo * n]) <= eps * (scale > 0 ? scale : 1.0)) {
                H[(lo - 1) + lo * n] = complex_zero();
                break;
            }
            lo--;
        }
        if (lo == hi) {  /* one eigenvalue deflated */
            hi--;
            its = 0;
            continue;
        }
        if (++its > cap) {
            converged = 0;
            break;
        }
        /* Wilkinson shift: the eigenvalue of the trailing 2 x 2 block nearer its last entry */
        const complex a = H[(hi - 1) + (hi - 1) * n], b = H[hi + (hi - 1) * n];
        const complex c = H[(hi - 1) + hi * n], d = H[hi + hi * n];
        const complex half_tr = complex_smul(0.5, complex_add(a, d));
        const complex disc = complex_sqrt(complex_add(complex_mul(complex_smul(0.5, complex_sub(a, d)),
                                                                  complex_smul(0.5, complex_sub(a, d))),
                                                      complex_mul(b, c)));
        const complex l1 = complex_add(half_tr, disc), l2 = complex_sub(half_tr, disc);
        complex mu = complex_mag(complex_sub(l1, d)) < complex_mag(complex_sub(l2, d)) ? l1 : l2;
        if (its % 11 == 10) mu = complex_add(d, cplx(cabs1(H[(hi - 1) + hi * n]), 0));  /* exceptional shift */
        /* implicit single-shift QR step on rows/columns lo..hi (bulge chase with Givens rotations) */
        complex x = complex_sub(H[lo + lo * n], mu), y = H[lo + (lo + 1) * n];
        for (int k = lo; k < hi; k++) {
            double cg;
            complex sg;
            givens(x, y, &cg, &sg);
            for (int col = (k > lo ? k - 1 : lo); col <= hi; col++) {  /* rows k, k+1 <- G (rows) */
                const complex t1 = H[col + k * n], t2 = H[col + (k + 1) * n];
                H[col + k * n] = complex_add(complex_smul(cg, t1), complex_mul(sg, t2));
                H[col + (k + 1) * n] = complex_sub(complex_smul(cg, t2), complex_mul(complex_conj(sg), t1));
            }
            const int rmax = k + 2 < hi ? k + 2 : hi;
            for (int row = lo; row <= rmax; row++) {  /* columns k, k+1 <- (columns) G^H */
                const complex t1 = H[k + row * n], t2 = H[(k + 1) + row * n];
                H[k + row * n] = complex_add(complex_smul(cg, t1), complex_mul(complex_conj(sg), t2));
                H[(k + 1) + row * n] = complex_sub(complex_smul(cg, t2), complex_mul(sg, t1));
            }
            if (k < hi - 1) {
                x = H[k + (k + 1) * n];
                y = H[k + (k + 2) * n];
            }
        }
    }
    for (int q = 0; q < n; q++) ev[q] = H[q + q * n];
    free(H);
    return converged;
}

matrix matrix_diagonalize(matrix A, double tol, int max_iterations) {
    (void)tol;
    need(A.i == A.j, "Matrix is not square for QL diagonalization!");
    const int n = A.i;
    matrix D = matrix_new(n, n);
    if (n == 0) return D;
    complex *ev = malloc((size_t)n * sizeof(complex));
    need(ev != NULL, "out of memory");
    schur_eigenvalues(A.m, n, max_iterations, ev);
    for (int q = 0; q < n; q++) D.m[q + q * n] = ev[q];
    free(ev);
    return D;
}

/* The two roots of lambda^2 - (a + d) lambda + (a d - b c) for [[a, b], [c, d]]:
 * (a + d +/- sqrt((a - d)^2 + 4 b c)) / 2, the root with the principal square root first. */
static void roots_2x2(complex a, complex b, complex c, complex d, complex *l1, complex *l2) {
    const complex amd = complex_sub(a, d);
    const complex root = complex_sqrt(complex_add(complex_mul(amd, amd), complex_smul(4, complex_mul(b, c))));
    *l1 = complex_smul(0.5, complex_add(complex_add(a, d), root));
    *l2 = complex_smul(0.5, complex_sub(complex_add(a, d), root));
}

/* Eigenvalues of any 2 x 2 matrix as diag(lambda1, lambda2) (matrix.c:1084). */
matrix matrix_diagonalize_2_by_2(matrix A) {
    need(A.i == 2 && A.j == 2, "Matrix is not 2x2 for 2x2 diagonalization!");
    matrix D = matrix_new(2, 2);
    roots_2x2(A.m[0], A.m[1], A.m[2], A.m[3], &D.m[0], &D.m[3]);
    return D;
}

static int by_magnitude(const void *x, const void *y) {
    const double mx = complex_mag(*(const complex *)x), my = complex_mag(*(const complex *)y);
    return (mx > my) - (mx < my);
}

/* Eigenvalues of a general square matrix on the diagonal of D (matrix.c:873).  The reference runs
 * unshifted QL sweeps until every entry above the diagonal but isolated first-superdiagonal ones
 * is below tol (at most max_iterations sweeps, then it prints a notice and stops), which leaves the
 * eigenvalues in increasing magnitude down the diagonal, and then re-solves the leading 2 x 2
 * block in closed form.  Here the eigenvalues come from the shifted Hessenberg QR above (converged
 * to working precision; tol unused), are placed in increasing magnitude (ties in deflation order),
 * and the leading pair is ordered as the closed form orders the roots of diag(d0, d1).  The
 * off-diagonal entries are zero; the reference's result keeps its lower triangle.  Equal
 * eigenvalues give finite results here (the reference's closed form takes sqrt(0) there). */
matrix matrix_diagonalize_non_hermitian(matrix A, double tol, int max_iterations) {
    (void)tol;
    need(A.i == A.j, "Matrix is not square for QL diagonalization!");
    const int n = A.i;
    matrix D = matrix_new(n, n);
    if (n == 0) return D;
    complex *ev = malloc((size_t)n * sizeof(complex));
    need(ev != NULL, "out of memory");
    if (!schur_eigenvalues(A.m, n, max_iterations, ev)) printf("QL algorithm reached maximum iterations!\n");
    /* insertion sort keeps equal magnitudes in deflation order (qsort is not stable) */
    for (int q = 1; q < n; q++) {
        const complex v = ev[q];
        int r = q - 1;
        while (r >= 0 && by_magnitude(&ev[r], &v) > 0) {
            ev[r + 1] = ev[r];
            r--;
        }
        ev[r + 1] = v;
    }
    /* the closed form on diag(d0, d1) puts first d0 + (principal sqrt((d0 - d1)^2) - (d0 - d1)) / 2,
     * i.e. d0 when d0 - d1 lies in the right half-plane (or on the positive imaginary axis), else
     * d1; decided here without the square root, which the series-based complex_sqrt returns as
     * NaN at 0 (equal eigenvalues) */
    if (n >= 2) {
        const complex dd = complex_sub(ev[0], ev[1]);
        if (!(dd.re > 0 || (dd.re == 0 && dd.im >= 0))) {
            const complex t = ev[0];
            ev[0] = ev[1];
            ev[1] = t;
        }
    }
    for (int q = 0; q < n; q++) D.m[q + q * n] = ev[q];
    free(ev);
    return D;
}

/* Eigenvector of A for an eigenvalue: inverse iteration on A - lambda I (the shift nudged by a
 * few ulps so the factorisation stays regular), normalised to unit 2-norm with the phase the
 * reference's construction gives (sum of the components real and positive, matrix.c:1114-1158). */
matrix matrix_eigenvector(matrix A, complex eigenvalue) {
    need(A.i == A.j, "A matrix is not square for eigenvector solve!");
    const int n = A.i;
    matrix x = matrix_new(n, 1);
    if (n == 0) return x;
    double an = 0.0;
    for (int q = 0; q < n * n; q++) an = fmax(an, complex_mag(A.m[q]));
    const double nudge = 1e-13 * (an > 0 ? an : 1.0);
    const complex lam = complex_add(eigenvalue, cplx(nudge, nudge));
    /* LU with partial pivoting of B = A - lam I, pivots floored at the nudge */
    complex *B = malloc((size_t)n * n * sizeof(complex));
    int *piv = malloc((size_t)n * sizeof(int));
    need(B && piv, "out of memory");
    memcpy(B, A.m, (size_t)n * n * sizeof(complex));
    for (int q = 0; q < n; q++) B[q + q * n] = complex_sub(B[q + q * n], lam);
    for (int col = 0; col < n; col++) {
        int p = col;
        for (int r = col + 1; r < n; r++)
            if (complex_mag(B[col + r * n]) > complex_mag(B[col + p * n])) p = r;
        piv[col] = p;
        if (p != col)
            for (int c = 0; c < n; c++) {
                const complex t = B[c + col * n];
                B[c + col * n] = B[c + p * n];
                B[c + p * n] = t;
            }
        if (complex_mag(B[col + col * n]) < nudge) B[col + col * n] = cplx(nudge, 0);
        const complex inv = complex_reciprocal(B[col + col * n]);
        for (int r = col + 1; r < n; r++) {
            const complex f = complex_mul(B[col + r * n], inv);
            B[col + r * n] = f;
            for (int c = col + 1; c < n; c++) B[c + r * n] = complex_sub(B[c + r * n], complex_mul(f, B[c + col * n]));
        }
    }
    for (int q = 0; q < n; q++) x.m[q] = cplx(1, 0);
    for (int it = 0; it < 3; it++) {
        for (int col = 0; col < n; col++)
            if (piv[col] != col) {
                const complex t = x.m[col];
                x.m[col] = x.m[piv[col]];
                x.m[piv[col]] = t;
            }
        for (int r = 0; r < n; r++)
            for (int c = 0; c < r; c++) x.m[r] = complex_sub(x.m[r], complex_mul(B[c + r * n], x.m[c]));
        for (int r = n - 1; r >= 0; r--) {
            complex s = x.m[r];
            for (int c = r + 1; c < n; c++) s = complex_sub(s, complex_mul(B[c + r * n], x.m[c]));
            x.m[r] = complex_mul(s, complex_reciprocal(B[r + r * n]));
        }
        matrix_vector_normalize_in_place(x);
    }
    complex sum = complex_zero();
    for (int q = 0; q < n; q++) sum = complex_add(sum, x.m[q]);
    const double ms = complex_mag(sum);
    if (ms > 0) matrix_complex_smul_in_place(complex_smul(1.0 / ms, complex_conj(sum)), x);
    free(B);
    free(piv);
    return x;
}
