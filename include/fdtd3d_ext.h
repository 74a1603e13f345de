/*
 * fdtd3d_ext.h -- extensions beyond the reference C surface (no counterpart in the reference).
 *
 * For callers that need more than the drop-in API: explicit host<->device synchronisation of
 * the timefield3d mirrors, timers on the grid's HIP stream, and the error text of the last
 * failing device call.  Every function returns 0 on success and a negative code on failure.
 */
#ifndef FDTD3D_EXT_H
#define FDTD3D_EXT_H

#include "timefield3d.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Upload ep/mu/sig, fields and sources from the host arrays of psi (after direct writes). */
int fdtd3d_ext_sync_to_device(timefield3d psi);
/* Refresh the host arrays fields, sources and fourier_fields of psi from the device. */
int fdtd3d_ext_sync_from_device(timefield3d psi);
/* Upload z plane k of the host sources array only (after direct writes to psi.sources). */
int fdtd3d_ext_upload_source_plane(timefield3d psi, int k);

/* Time the work enqueued on psi's stream between start and stop (HIP events on that stream). */
int fdtd3d_ext_timer_start(timefield3d psi);
int fdtd3d_ext_timer_stop(timefield3d psi, double *elapsed_ms);
int fdtd3d_ext_synchronize(timefield3d psi);

/* Run n reference steps (timefield3d_apply_timestep_omp_full semantics) without host work in
 * between; sources stay as last set. */
int fdtd3d_ext_steps(timefield3d psi, double delta_t, int n);

const char *fdtd3d_ext_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
