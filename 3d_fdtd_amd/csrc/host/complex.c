/* Scalar complex arithmetic of the drop-in surface (complex.h).  Semantics follow complex.c of
 * the reference, including its truncated pi (3.14159265359, complex.c:3) for the angle range
 * and the power-series exponential/cosine/sine summed until a term drops below 1e-30
 * (complex.c:169-254); callers rely on the exact twiddle values these produce. */
#include "complex.h"

#define CPX_PI 3.14159265359
#define CPX_SERIES_TOL 1e-30

double complex_mag(complex x) { return sqrt(x.re * x.re + x.im * x.im); }

double complex_angle(complex x) {
    double a = atan(x.im / x.re);
    if (x.re < 0) a += CPX_PI;
    else if (x.im < 0) a += 2 * CPX_PI;
    return a;
}

double complex_real(complex x) { return x.re; }
double complex_imag(complex x) { return x.im; }

complex complex_zero() {
    complex z = {0.0, 0.0};
    return z;
}

complex complex_from_doubles(double real, double imag) {
    complex z = {real, imag};
    return z;
}

complex complex_conj(complex x) {
    complex z = {x.re, -x.im};
    return z;
}

complex complex_add(complex x, complex y) {
    complex z = {x.re + y.re, x.im + y.im};
    return z;
}

complex complex_sub(complex x, complex y) {
    complex z = {x.re - y.re, x.im - y.im};
    return z;
}

complex complex_smul(double s, complex x) {
    complex z = {s * x.re, s * x.im};
    return z;
}

complex complex_mul(complex x, complex y) {
    complex z = {x.re * y.re - x.im * y.im, x.re * y.im + x.im * y.re};
    return z;
}

/* x * conj(y) / |y|^2 with |y| formed as a root and squared again (complex.c:125-139). */
complex complex_div(complex x, complex y) {
    complex num = complex_mul(x, complex_conj(y));
    double m = complex_mag(y);
    double inv = 1.0 / (m * m);
    return complex_smul(inv, num);
}

complex complex_reciprocal(complex x) { return complex_div(complex_from_doubles(1, 0), x); }

complex complex_sqrt(complex x) {
    double mag = sqrt(complex_mag(x));
    double ang = complex_angle(x);
    if (ang > CPX_PI) ang -= 2 * CPX_PI;
    ang /= 2;
    return complex_smul(mag, complex_exp(complex_from_doubles(0, ang)));
}

/* sum_n x^n / n! until |x^n/n!| < 1e-30 */
complex complex_exp(complex x) {
    complex term = {1.0, 0.0}, acc = {0.0, 0.0}, pw = {1.0, 0.0};
    double fact = 1.0;
    int n = 0;
    while (complex_mag(term) >= CPX_SERIES_TOL) {
        double inv = 1.0 / fact;
        term = complex_smul(inv, pw);
        acc = complex_add(acc, complex_smul(inv, pw));
        n++;
        fact *= n;
        pw = complex_mul(pw, x);
    }
    return acc;
}

/* even (cos) or odd (sin) terms of the same series with alternating sign */
static complex series_trig(complex x, int parity) {
    complex term = {1.0, 0.0}, acc = {0.0, 0.0}, pw = {1.0, 0.0};
    double fact = 1.0;
    int n = 0;
    while (complex_mag(term) >= CPX_SERIES_TOL) {
        if (n % 2 == parity) {
            double inv = 1.0 / fact;
            term = complex_smul(inv, pw);
            acc = complex_add(acc, complex_smul(inv, pw));
            pw = complex_smul(-1, pw);
        }
        n++;
        fact *= n;
        pw = complex_mul(pw, x);
    }
    return acc;
}

complex complex_cos(complex x) { return series_trig(x, 0); }
complex complex_sin(complex x) { return series_trig(x, 1); }
