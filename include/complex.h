/*
 * complex.h -- scalar complex type and arithmetic of the drop-in C surface.
 *
 * Replaces the reference interface complex.h:11-38 one for one (same type layout, names and
 * signatures) so that drivers written against the reference compile and link unchanged.
 * Semantics follow complex.c (angle in [0, 2pi), Taylor-series exp/cos/sin).
 */
#ifndef _complex_h
#define _complex_h

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

/* {re, im} pair of doubles, passed and returned by value (complex.h:11-15). */
struct complex_ {
    double re, im;
};
typedef struct complex_ complex;

double complex_mag(complex x);                 /* |x|                      (complex.c:6)   */
double complex_angle(complex x);               /* arg x in [0, 2pi)        (complex.c:15)  */
double complex_real(complex x);                /*                          (complex.c:37)  */
double complex_imag(complex x);                /*                          (complex.c:46)  */
complex complex_zero();                        /*                          (complex.c:55)  */
complex complex_from_doubles(double real, double imag); /*                 (complex.c:65)  */
complex complex_conj(complex x);               /*                          (complex.c:75)  */
complex complex_add(complex x, complex y);     /*                          (complex.c:85)  */
complex complex_sub(complex x, complex y);     /*                          (complex.c:95)  */
complex complex_smul(double scalar, complex x);/*                          (complex.c:105) */
complex complex_mul(complex x, complex y);     /*                          (complex.c:115) */
complex complex_div(complex x, complex y);     /*                          (complex.c:125) */
complex complex_reciprocal(complex x);         /*                          (complex.c:141) */
complex complex_sqrt(complex x);               /* principal root           (complex.c:151) */
complex complex_exp(complex x);                /* power series to 1e-30    (complex.c:169) */
complex complex_cos(complex x);                /*                          (complex.c:195) */
complex complex_sin(complex x);                /*                          (complex.c:225) */

#ifdef __cplusplus
}
#endif
#endif
