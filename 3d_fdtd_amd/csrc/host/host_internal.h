/* Declarations shared by the host translation units (not part of the public surface). */
#ifndef FDTD3D_HOST_INTERNAL_H
#define FDTD3D_HOST_INTERNAL_H

#include "fdtd3d_dev.h"
#include "timefield3d.h"

struct ext_state;

typedef struct {
    const double *key;  /* host fields pointer of the grid */
    fdtd_dev *dev;
    int mats_dirty;
    struct ext_state *ext;  /* boundary / run configuration of the extension API, or NULL */
} tf_record;

tf_record *tf_lookup(timefield3d psi);
fdtd_dev *tf_dev(timefield3d psi);
void dev_check(int rc, const char *where);
void tf_twiddle(double phase, double *re, double *im);

/* One step of timefield3d_apply_timestep_omp_full semantics under the grid's boundary mode. */
void tf_step(timefield3d psi, double dt);
void ext_release(tf_record *r);

#endif
