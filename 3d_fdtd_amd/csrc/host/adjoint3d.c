/* Host side of the adjoint design surface (adjoint3d.h).  Per-step field stores gather into
 * device histories (fdtd_dev_taper_store / fdtd_dev_block_store, keyed by the struct's history
 * pointers); gradients are reduced on the device and written to the host arrays the reference
 * exposes.  The design updates are scalar host logic restated from adjoint3d.c:420-664. */
#include "adjoint3d.h"

#include <string.h>

#include "host_internal.h"

/* ------------------------------------------------------------------------------------------ */
/* 3-D block adjoint (adjoint3d.c:6-185)                                                      */
/* ------------------------------------------------------------------------------------------ */

static void box_of(adjoint3d b, int box[6]) {
    box[0] = b.x_start;
    box[1] = b.x_end;
    box[2] = b.y_start;
    box[3] = b.y_end;
    box[4] = b.z_start;
    box[5] = b.z_end;
}

static long long block_cells(adjoint3d b) {
    return (long long)(b.x_end - b.x_start) * (b.y_end - b.y_start) * (b.z_end - b.z_start);
}

/* The forward history lives on the device; the host forward_fields array is a 1-element
 * placeholder that only serves as the history's key. */
adjoint3d ajoint3d_new(int x_start, int x_end, int y_start, int y_end, int z_start, int z_end, int t) {
    adjoint3d b;
    b.t = t;
    b.x_start = x_start;
    b.x_end = x_end;
    b.y_start = y_start;
    b.y_end = y_end;
    b.z_start = z_start;
    b.z_end = z_end;
    b.forward_fields = calloc(1, sizeof(double));
    long long n = block_cells(b);
    b.gradient = calloc(n > 0 ? n : 1, sizeof(double));
    return b;
}

void adjoint3d_free(adjoint3d block) {
    fdtd_dev_block_release(block.forward_fields);
    free(block.forward_fields);
    free(block.gradient);
}

void adjointd3d_free(adjoint3d block) { adjoint3d_free(block); }

void adjoint3d_store_fwd_fields(adjoint3d block, timefield3d f, int timestep) {
    int box[6];
    box_of(block, box);
    dev_check(fdtd_dev_block_store(tf_dev(f), block.forward_fields, box, block.t, timestep), "store_fwd_fields");
}

/* gradient += sum_m fwd(timestep) * E(now), then the host gradient array is refreshed. */
void adjoint3d_update_gradient(adjoint3d block, timefield3d f, int timestep) {
    int box[6];
    box_of(block, box);
    dev_check(fdtd_dev_block_update_gradient(tf_dev(f), block.forward_fields, box, block.t, timestep, block.gradient),
              "update_gradient");
}

void adjoint3d_reset_gradient(adjoint3d block) {
    memset(block.gradient, 0, sizeof(double) * block_cells(block));
    dev_check(fdtd_dev_block_reset(block.forward_fields, 0), "reset_gradient");
}

void adjoint3d_reset_forward_fields(adjoint3d block) {
    dev_check(fdtd_dev_block_reset(block.forward_fields, 1), "reset_forward_fields");
}

/* Continuous permittivity step along the normalised gradient, clamped to [ep_min, ep_max]
 * (adjoint3d.c:131-185). */
void adjoint3d_adjust_permittivity(adjoint3d block, timefield3d f, double omega, double improvement, double ep_min,
                                   double ep_max) {
    int nx = block.x_end - block.x_start, ny = block.y_end - block.y_start, nz = block.z_end - block.z_start;
    long long n = (long long)nx * ny * nz;
    double mag = 0;
    for (long long q = 0; q < n; q++) mag += block.gradient[q] * block.gradient[q];
    mag = sqrt(mag) * f.disc * f.disc;
    double *dir = malloc(sizeof(double) * (n > 0 ? n : 1));
    for (long long q = 0; q < n; q++) dir[q] = block.gradient[q] * f.disc * f.disc / mag;
    mag *= omega;
    for (int i = block.x_start; i < block.x_end; i++)
        for (int j = block.y_start; j < block.y_end; j++)
            for (int k = block.z_start; k < block.z_end; k++) {
                long long g = (i - block.x_start) + (long long)nx * ((j - block.y_start) + (long long)ny * (k - block.z_start));
                long long e = 3 * (i + (long long)f.i * (j + (long long)f.j * k));
                for (int m = 0; m < 3; m++) {
                    f.ep[e + m] += (improvement / mag) * dir[g];
                    if (f.ep[e + m] < ep_min) f.ep[e + m] = ep_min;
                    if (f.ep[e + m] > ep_max) f.ep[e + m] = ep_max;
                }
            }
    free(dir);
    tf_lookup(f)->mats_dirty = 1;
}

/* ------------------------------------------------------------------------------------------ */
/* taper (adjoint3d.c:187-664)                                                                */
/* ------------------------------------------------------------------------------------------ */

/* Histories live on the device; forward_fields / adjoint_fields are 1-element placeholders
 * used as keys. */
adjoint3d_taper adjoint3d_taper_new(double ep, double mu, int seg_num, int start, int end, int *left, int *right,
                                    int bottom, int top, int duration) {
    adjoint3d_taper t;
    t.ep = ep;
    t.mu = mu;
    t.seg_num = seg_num;
    t.start = start;
    t.end = end;
    t.bottom = bottom;
    t.top = top;
    t.duration = duration;
    t.left = malloc(sizeof(int) * (seg_num + 1));
    t.right = malloc(sizeof(int) * (seg_num + 1));
    memcpy(t.left, left, sizeof(int) * (seg_num + 1));
    memcpy(t.right, right, sizeof(int) * (seg_num + 1));
    t.forward_fields = calloc(1, sizeof(double));
    t.adjoint_fields = calloc(1, sizeof(double));
    int n = 2 * (top - bottom) * (end - start);
    t.ep_gradient = calloc(n > 0 ? n : 1, sizeof(double));
    t.left_grad = calloc(seg_num + 1, sizeof(double));
    t.right_grad = calloc(seg_num + 1, sizeof(double));
    return t;
}

void adjoint3d_taper_insert(timefield3d psi, adjoint3d_taper t) {
    int L = t.end - t.start;
    for (int n = 0; n < t.seg_num; n++) {
        int s0 = t.start + (n * L) / t.seg_num;
        int s1 = t.start + ((n + 1) * L) / t.seg_num;
        timefield3d_set_taper(psi, t.ep, t.mu, t.left[n], t.right[n], t.left[n + 1], t.right[n + 1], s0, s1, t.bottom,
                              t.top);
    }
}

static void taper_store(adjoint3d_taper t, timefield3d f, int timestep, int which) {
    dev_check(fdtd_dev_taper_store(tf_dev(f), t.forward_fields, which, t.seg_num, t.start, t.end, t.left, t.right,
                                   t.bottom, t.top, t.duration, timestep),
              "taper store");
}

void adjoint3d_taper_store_forward_fields(adjoint3d_taper t, timefield3d f, int timestep) {
    taper_store(t, f, timestep, 0);
}

void adjoint3d_taper_store_adjoint_fields(adjoint3d_taper t, timefield3d f, int timestep) {
    taper_store(t, f, timestep, 1);
}

void adjoint3d_taper_calculate_ep_gradient(adjoint3d_taper t) {
    dev_check(fdtd_dev_taper_ep_gradient(t.forward_fields, t.top - t.bottom, t.end - t.start, t.duration,
                                         t.ep_gradient),
              "calculate_ep_gradient");
}

/* Hat-weighted projection of the edge gradient onto interior control points, left side
 * negated, end points zero (adjoint3d.c:375-418). */
void adjoint3d_taper_calculate_gradients_from_ep_gradient(adjoint3d_taper t, double ep_diff) {
    int L = t.end - t.start, H = t.top - t.bottom;
    t.left_grad[0] = 0;
    t.right_grad[0] = 0;
    for (int n = 1; n < t.seg_num; n++) {
        t.left_grad[n] = 0;
        t.right_grad[n] = 0;
        int sb = ((n - 1) * L) / t.seg_num, sm = (n * L) / t.seg_num, sf = ((n + 1) * L) / t.seg_num;
        for (int k = sb; k < sm; k++)
            for (int j = 0; j < H; j++) {
                double w = (double)(k - sb) / (double)(sm - sb);
                t.right_grad[n] += w * t.ep_gradient[0 + 2 * (j + H * k)] * ep_diff;
                t.left_grad[n] -= w * t.ep_gradient[1 + 2 * (j + H * k)] * ep_diff;
            }
        for (int k = sm; k < sf; k++)
            for (int j = 0; j < H; j++) {
                double w = (double)(sf - k) / (double)(sf - sm);
                t.right_grad[n] += w * t.ep_gradient[0 + 2 * (j + H * k)] * ep_diff;
                t.left_grad[n] -= w * t.ep_gradient[1 + 2 * (j + H * k)] * ep_diff;
            }
    }
    t.left_grad[t.seg_num] = 0;
    t.right_grad[t.seg_num] = 0;
}

/* The same projection from a 3-D block gradient, sampling edge pixels with integer division
 * (adjoint3d.c:420-469). */
void adjoint3d_taper_calculate_gradients(adjoint3d_taper t, adjoint3d b, double ep_diff) {
    int nx = b.x_end - b.x_start, ny = b.y_end - b.y_start;
    int L = t.end - t.start;
    t.left_grad[0] = 0;
    t.right_grad[0] = 0;
    for (int n = 1; n < t.seg_num; n++) {
        t.left_grad[n] = 0;
        t.right_grad[n] = 0;
        int sb = t.start + ((n - 1) * L) / t.seg_num, sm = t.start + (n * L) / t.seg_num;
        int sf = t.start + ((n + 1) * L) / t.seg_num;
        for (int k = sb; k < sm; k++)
            for (int j = t.bottom; j < t.top; j++) {
                double w = (double)(k - sb) / (double)(sm - sb);
                int i = t.right[n - 1] + ((t.right[n] - t.right[n - 1]) * (k - sb)) / (sm - sb);
                long long g = (i - b.x_start) + (long long)nx * ((j - b.y_start) + (long long)ny * (k - b.z_start));
                t.right_grad[n] += w * b.gradient[g] * ep_diff;
                i = t.left[n - 1] + ((t.left[n] - t.left[n - 1]) * (k - sb)) / (sm - sb);
                g = (i - b.x_start) + (long long)nx * ((j - b.y_start) + (long long)ny * (k - b.z_start));
                t.left_grad[n] -= w * b.gradient[g] * ep_diff;
            }
        for (int k = sm; k < sf; k++)
            for (int j = t.bottom; j < t.top; j++) {
                double w = (double)(sf - k) / (double)(sf - sm);
                int i = t.right[n] + ((t.right[n + 1] - t.right[n]) * (k - sm)) / (sf - sm);
                long long g = (i - b.x_start) + (long long)nx * ((j - b.y_start) + (long long)ny * (k - b.z_start));
                t.right_grad[n] += w * b.gradient[g] * ep_diff;
                i = t.left[n] + ((t.left[n + 1] - t.left[n]) * (k - sm)) / (sf - sm);
                g = (i - b.x_start) + (long long)nx * ((j - b.y_start) + (long long)ny * (k - b.z_start));
                t.left_grad[n] -= w * b.gradient[g] * ep_diff;
            }
    }
    t.left_grad[t.seg_num] = 0;
    t.right_grad[t.seg_num] = 0;
}

static double sym_direction(adjoint3d_taper t, double disc, double *dir) {
    double mag = 0;
    for (int i = 0; i < t.seg_num + 1; i++) mag += (t.right_grad[i] - t.left_grad[i]) * (t.right_grad[i] - t.left_grad[i]);
    mag = sqrt(mag) * 0.5 * disc * disc;
    for (int i = 0; i < t.seg_num + 1; i++) dir[i] = (t.right_grad[i] - t.left_grad[i]) * 0.5 * disc * disc / mag;
    return mag;
}

/* Symmetric width update: both edges move by round(goal/|g| * dir); while every move is below
 * half a pixel the smallest non-zero component is zeroed (adjoint3d.c:471-562). */
int adjoint3d_taper_update_symmetric(adjoint3d_taper t, double goal, double disc, int min_width) {
    double *dir = malloc(sizeof(double) * (t.seg_num + 1));
    double mag = sym_direction(t, disc, dir);
    int unchanged = 1;
    for (int i = 1; i < t.seg_num; i++)
        if (fabs((goal / mag) * dir[i]) >= 0.5) unchanged = 0;
    int zeroed = 0;
    while (unchanged == 1 && zeroed < t.seg_num - 1) {
        double best = 1000;
        int at = 0;
        for (int i = 1; i < t.seg_num; i++) {
            double a = fabs(dir[i]) * 2 * disc;
            if (a < best && a > 0) { best = a; at = i; }
        }
        t.right_grad[at] = 0;
        t.left_grad[at] = 0;
        mag = sym_direction(t, disc, dir);
        unchanged = 1;
        for (int i = 1; i < t.seg_num; i++)
            if (fabs((goal / mag) * dir[i]) >= 0.5) unchanged = 0;
        printf("zeroed out index %i\n", at);
        zeroed++;
    }
    for (int i = 1; i < t.seg_num; i++) {
        t.left[i] -= (int)(round((goal / mag) * dir[i]));
        t.right[i] += (int)(round((goal / mag) * dir[i]));
        if (t.right[i] - t.left[i] < min_width) {
            t.right[i] = (t.right[i] + t.left[i] + min_width) / 2;
            t.left[i] = (t.right[i] + t.left[i] - min_width) / 2;
        }
    }
    free(dir);
    return unchanged;
}

static double asym_direction(adjoint3d_taper t, double disc, double *dir) {
    double mag = 0;
    for (int i = 0; i < t.seg_num + 1; i++) {
        mag += t.right_grad[i] * t.right_grad[i];
        mag += t.left_grad[i] * t.left_grad[i];
    }
    mag = sqrt(mag) * 0.5 * disc * disc;
    for (int i = 0; i < t.seg_num + 1; i++) {
        dir[2 * i] = t.right_grad[i] * 0.5 * disc * disc / mag;
        dir[2 * i + 1] = t.left_grad[i] * 0.5 * disc * disc / mag;
    }
    return mag;
}

/* Independent edge moves (adjoint3d.c:564-664). */
int adjoint3d_taper_update_asymmetric(adjoint3d_taper t, double goal, double disc, int min_width) {
    double *dir = malloc(sizeof(double) * 2 * (t.seg_num + 1));
    double mag = asym_direction(t, disc, dir);
    int unchanged = 1;
    for (int i = 2; i < 2 * t.seg_num; i++)
        if (round(fabs((goal / mag) * dir[i])) >= 1) unchanged = 0;
    int zeroed = 0;
    while (unchanged == 1 && zeroed < 2 * (t.seg_num - 2)) {
        double best = 1e10;
        int at = 0;
        for (int i = 2; i < 2 * t.seg_num; i++) {
            double a = fabs(dir[i]);
            if (a < best && a > 0) { best = a; at = i; }
        }
        if (at % 2 == 0) t.right_grad[at / 2] = 0;
        else t.left_grad[at / 2] = 0;
        mag = asym_direction(t, disc, dir);
        unchanged = 1;
        for (int i = 2; i < 2 * t.seg_num; i++)
            if (round(fabs((goal / mag) * dir[i])) >= 1) unchanged = 0;
        printf("zeroed out index %i\n", at);
        zeroed++;
    }
    for (int i = 1; i < t.seg_num; i++) {
        t.right[i] += (int)(round((goal / mag) * dir[2 * i]));
        t.left[i] += (int)(round((goal / mag) * dir[2 * i + 1]));
        if (t.right[i] - t.left[i] < min_width) {
            t.right[i] = (t.right[i] + t.left[i] + min_width) / 2;
            t.left[i] = (t.right[i] + t.left[i] - min_width) / 2;
        }
    }
    free(dir);
    return unchanged;
}
