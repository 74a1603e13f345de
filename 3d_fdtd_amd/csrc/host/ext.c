/* Extension API (fdtd3d_ext.h): host<->device synchronisation, timers, multi-step runs, and the
 * step dispatch used by timefield3d_apply_timestep_omp_full. */
#include "fdtd3d_ext.h"

#include <string.h>

#include "fdtd3d_dev.h"
#include "host_internal.h"

struct ext_state {
    int boundary; /* reserved for boundary kinds beyond the reference's Mur faces */
};

void ext_release(tf_record *r) {
    free(r->ext);
    r->ext = NULL;
}

void tf_step(timefield3d psi, double dt) { dev_check(fdtd_dev_step(tf_dev(psi), dt, 0), "apply_timestep_omp_full"); }

const char *fdtd3d_ext_last_error(void) { return fdtd_dev_last_error(); }

int fdtd3d_ext_sync_to_device(timefield3d psi) {
    tf_lookup(psi)->mats_dirty = 1;
    fdtd_dev *d = tf_dev(psi); /* creates the device side if needed and uploads materials */
    int rc = fdtd_dev_set_fields(d, psi.fields);
    if (rc) return rc;
    return fdtd_dev_set_sources(d, psi.sources);
}

int fdtd3d_ext_sync_from_device(timefield3d psi) {
    fdtd_dev *d = tf_dev(psi);
    int rc = fdtd_dev_get_fields(d, psi.fields);
    if (rc) return rc;
    rc = fdtd_dev_get_sources(d, psi.sources);
    if (rc) return rc;
    return fdtd_dev_get_fourier(d, (double *)psi.fourier_fields);
}

int fdtd3d_ext_upload_source_plane(timefield3d psi, int k) {
    if (k < 0 || k >= psi.k) return -2;
    return fdtd_dev_set_source_plane(tf_dev(psi), k, psi.sources + 6LL * psi.i * psi.j * k, NULL);
}

int fdtd3d_ext_timer_start(timefield3d psi) { return fdtd_dev_timer_start(tf_dev(psi)); }

int fdtd3d_ext_timer_stop(timefield3d psi, double *elapsed_ms) {
    return fdtd_dev_timer_stop(tf_dev(psi), elapsed_ms);
}

int fdtd3d_ext_synchronize(timefield3d psi) { return fdtd_dev_synchronize(tf_dev(psi)); }

int fdtd3d_ext_steps(timefield3d psi, double dt, int n) {
    for (int s = 0; s < n; s++) tf_step(psi, dt);
    return 0;
}
