/*
 * adjoint3d.h -- adjoint design surface of the drop-in C API (MI355X implementation).
 *
 * Replaces the reference interface adjoint3d.h:16-69 one for one (struct layouts, names,
 * signatures, including the reference's spelling ajoint3d_new).  The per-step field stores
 * gather into device-resident histories (FP32, the precision the device fields carry) keyed
 * by the struct's history pointers; the gradient reductions run on the device in FP64 and write
 * the host arrays the reference exposes (ep_gradient, left_grad, right_grad, gradient).  The
 * host history arrays (forward_fields, adjoint_fields) are not mirrored.
 */
#ifndef _adjoint3d_h
#define _adjoint3d_h

#include "complex.h"
#include "matrix.h"
#include "mode.h"
#include "timefield3d.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 3-D block (density) adjoint, adjoint3d.h:16-34 */
struct adjoint3d_ {
    int t;
    int x_start, x_end, y_start, y_end, z_start, z_end;
    double *forward_fields, *gradient;
};
typedef struct adjoint3d_ adjoint3d;

adjoint3d ajoint3d_new(int x_start, int x_end, int y_start, int y_end, int z_start, int z_end, int t);
void adjoint3d_free(adjoint3d block);
void adjointd3d_free(adjoint3d block); /* the name the reference source defines (adjoint3d.c:42) */
void adjoint3d_store_fwd_fields(adjoint3d block, timefield3d main_fields, int timestep);
void adjoint3d_update_gradient(adjoint3d block, timefield3d main_fields, int timestep);
void adjoint3d_reset_gradient(adjoint3d block);
void adjoint3d_reset_forward_fields(adjoint3d block);
void adjoint3d_adjust_permittivity(adjoint3d block, timefield3d main_fields, double omega, double improvement,
                                   double ep_min, double ep_max);

/* piecewise-linear taper, adjoint3d.h:37-55 */
struct adjoint3d_taper_ {
    double ep;
    double mu;
    int seg_num;
    int start;
    int end;
    int bottom;
    int top;
    int duration;
    int *left;
    int *right;
    double *forward_fields;
    double *adjoint_fields;
    double *ep_gradient;
    double *left_grad;
    double *right_grad;
};
typedef struct adjoint3d_taper_ adjoint3d_taper;

adjoint3d_taper adjoint3d_taper_new(double ep, double mu, int seg_num, int start, int end, int *left, int *right,
                                    int bottom, int top, int duration);
void adjoint3d_taper_insert(timefield3d psi, adjoint3d_taper taper);
void adjoint3d_taper_store_forward_fields(adjoint3d_taper taper, timefield3d main_fields, int timestep);
void adjoint3d_taper_store_adjoint_fields(adjoint3d_taper taper, timefield3d main_fields, int timestep);
void adjoint3d_taper_calculate_ep_gradient(adjoint3d_taper taper);
void adjoint3d_taper_calculate_gradients_from_ep_gradient(adjoint3d_taper taper, double ep_diff);
void adjoint3d_taper_calculate_gradients(adjoint3d_taper taper, adjoint3d block, double ep_diff);
int adjoint3d_taper_update_symmetric(adjoint3d_taper taper, double improvement_goal, double disc, int min_width);
int adjoint3d_taper_update_asymmetric(adjoint3d_taper taper, double improvement_goal, double disc, int min_width);

#ifdef __cplusplus
}
#endif
#endif
