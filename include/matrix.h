/*
 * matrix.h -- dense complex matrix type of the drop-in C surface.
 *
 * Replaces the reference interface matrix.h:12-59 (same struct layout, names, signatures).
 * Only the mode data path uses it on the FDTD hot path (mode.fields is a matrix); the dense
 * linear algebra serves the host-side mode solver.
 */
#ifndef _matrix_h
#define _matrix_h

#include "complex.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

/* i rows x j columns, row-major complex entries (matrix.h:12-16). */
struct matrix_ {
    int i, j;
    complex *m;
};
typedef struct matrix_ matrix;

matrix matrix_new(int i, int j);                       /* zero-filled        (matrix.c:4)    */
void matrix_free(matrix A);
void matrix_print(matrix A);
matrix matrix_copy(matrix A);
void matrix_copy_in_place(matrix A, matrix B);
matrix matrix_get_column(matrix A, int j);
matrix matrix_get_row(matrix A, int i);
void matrix_set_column(matrix A, matrix col, int j);
void matrix_set_row(matrix A, matrix row, int i);
matrix matrix_get_submatrix(matrix A, int i, int j);   /* row i and column j removed (matrix.c:162) */
complex matrix_determinant(matrix A);                  /* square A (matrix.c:200)                    */
complex matrix_trace(matrix A);
matrix matrix_hermitian(matrix A);
matrix matrix_smul(double scalar, matrix A);
matrix matrix_complex_smul(complex scalar, matrix A);
matrix matrix_add(matrix A, matrix B);
matrix matrix_sub(matrix A, matrix B);
void matrix_smul_in_place(double scalar, matrix A);
void matrix_complex_smul_in_place(complex scalar, matrix A);
void matrix_add_in_place(matrix A, matrix B);
void matrix_sub_in_place(matrix A, matrix B);
matrix matrix_mul(matrix A, matrix B);
matrix matrix_linear_solve(matrix A, matrix B);        /* Gaussian elimination, partial pivot */
matrix matrix_inverse(matrix A);
void matrix_gram_schmidt(matrix A);
/* Eigenvalues of a square matrix as a diagonal matrix (matrix.c:672: there by a QL iteration
 * stopped at an off-diagonal ratio tol; here converged to working precision) and the unit
 * eigenvector for one of them (matrix.c:1114; phase: components sum to a positive real). */
matrix matrix_diagonalize(matrix A, double tol, int max_iterations);
/* Eigenvalues of a general square matrix on the diagonal, in increasing magnitude with the leading
 * pair in closed-form root order (matrix.c:873; off-diagonal entries zero here), and the two
 * roots of any 2 x 2 matrix as diag(l1, l2), l1 with the principal square root (matrix.c:1084). */
matrix matrix_diagonalize_non_hermitian(matrix A, double tol, int max_iterations);
matrix matrix_diagonalize_2_by_2(matrix A);
matrix matrix_eigenvector(matrix A, complex eigenvalue);

complex matrix_vector_inner_product(matrix A, matrix B);
double matrix_vector_norm(matrix A);
matrix matrix_vector_normalize(matrix A);
void matrix_vector_normalize_in_place(matrix A);

#ifdef __cplusplus
}
#endif
#endif
