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
