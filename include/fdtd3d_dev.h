/*
 * fdtd3d_dev.h -- the thin C-ABI between the C host layer and the HIP (gfx950) kernels.
 *
 * Plain C types only (no HIP or torch types): pointers, sizes, doubles.  Every call enqueues on
 * the grid's own HIP stream; calls that return data to the host synchronise that stream.  All
 * functions return 0 on success and a negative code on failure, with a message available from
 * fdtd_dev_last_error().  Array arguments use the reference's AoS layout (timefield3d.h:15-27):
 * cell c = i + I*(j + J*k), six components per cell for fields/sig/sources, three for ep/mu.
 *
 * Each entry point replaces the compute inside one reference function; the citation names it.
 */
#ifndef FDTD3D_DEV_H
#define FDTD3D_DEV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fdtd_dev fdtd_dev;

const char *fdtd_dev_last_error(void);

/* Device state for an I x J x K grid (timefield3d.c:6-41). */
fdtd_dev *fdtd_dev_create(int I, int J, int K, double disc);
void fdtd_dev_destroy(fdtd_dev *d);                                           /* timefield3d.c:43  */

/* Materials from host AoS float arrays: unique (ep[3], mu[3], sig[6]) tuples become a table and
 * each cell stores a 1-byte id (2-byte above 256 tuples, 4-byte above 65536: a per-cell graded
 * medium such as a block after adjoint3d_adjust_permittivity, whose table is then as long as the
 * grid).  Grids with 4-byte ids always step on the split kernels; batched runs and the fused and
 * leap launch shapes refuse them.  mu or sig may be NULL: mu = 1 and sig = 0 everywhere. */
int fdtd_dev_set_materials(fdtd_dev *d, const float *ep, const float *mu, const float *sig);

/* Extruded geometry rasterised on the device (no reference counterpart; it serves the design
 * loops, which rebuild a grid per candidate): a table of nmat materials (ep 3 per entry, mu 3 and
 * sig 6 per entry or NULL for 1 / 0) and, for the rows j0 <= j < j1, one id per (k, i) in ids[k * I
 * + i]; every other cell takes `background`.  This is the structure of the reference's waveguide
 * and taper setters (timefield3d.c:1458-1535) and of adjoint3d_taper_insert: the host uploads
 * K * I ids instead of 3 * N permittivities.  The grid keeps the K * I table on the device while
 * its id volume equals it: the split updates of CPML grids and batched runs of such grids then
 * read each cell's id from the table, not from the volume.  A later fdtd_dev_set_materials or
 * paint drops the table. */
int fdtd_dev_set_materials_extruded(fdtd_dev *d, int nmat, const float *ep, const float *mu, const float *sig,
                                    const uint16_t *ids, int j0, int j1, int background);

/* The reference's geometry setters applied to the device material ids, for grids whose device
 * materials are current (no per-cell upload; the host builds a table remap and a kernel rewrites
 * the ids of the region).  Each cell of the region keeps its conductivities.
 * paint_taper: the rule of timefield3d_set_taper (timefield3d.c:1511-1535): for y in [bottom, top)
 * and z in [start_z, end_z), x in [start_left + round(sl (z - start_z)), start_right +
 * round(sr (z - start_z))) with the slopes sl, sr in double; ep and mu of all three components.
 * paint_prism: x in [x0, x1), y in [y0, y1), z in [z0, z1) (timefield3d.c:1455-1474, the
 * waveguide setter of 1506-1509); ep, and mu when set_mu.  Cells outside the grid are skipped. */
int fdtd_dev_paint_taper(fdtd_dev *d, float ep, float mu, int start_left, int start_right, int end_left,
                         int end_right, int start_z, int end_z, int bottom, int top);
int fdtd_dev_paint_prism(fdtd_dev *d, float ep, float mu, int set_mu, int x0, int x1, int y0, int y1, int z0, int z1);
/* Per-cell materials as the device holds them (AoS ep[3N], mu[3N], sig[6N]; NULL skips one), and
 * counts[0] full material uploads, counts[1] device repaints, counts[2] table entries. */
int fdtd_dev_get_materials(fdtd_dev *d, float *ep, float *mu, float *sig);
int fdtd_dev_material_counts(fdtd_dev *d, long long counts[3]);

/* Fields / sources / shells / DFT accumulators between host AoS arrays and the device. */
int fdtd_dev_set_fields(fdtd_dev *d, const double *fields);
int fdtd_dev_get_fields(fdtd_dev *d, double *fields);                          /* timefield3d.c:123 */
int fdtd_dev_set_sources(fdtd_dev *d, const float *sources);
int fdtd_dev_get_sources(fdtd_dev *d, float *sources);
int fdtd_dev_zero(fdtd_dev *d);                                                /* timefield3d.c:108 */
int fdtd_dev_get_fourier(fdtd_dev *d, double *fourier_interleaved);
int fdtd_dev_set_fourier(fdtd_dev *d, const double *fourier_interleaved);

/* Time stepping.  scheme 0 = semi-implicit conductivity (timefield3d.c:1679-1837),
 * 1 = explicit (timefield3d.c:1537-1677).  [k0, k1) bounds the z range like the reference's
 * _zsection variants; interior cells are [1, N-2] in every axis. */
int fdtd_dev_snapshot_shells(fdtd_dev *d);                                     /* timefield3d.c:246 */
int fdtd_dev_set_shells(fdtd_dev *d, const double *old_fields);
int fdtd_dev_update_E(fdtd_dev *d, double dt, int k0, int k1, int scheme);    /* timefield3d.c:1759 */
int fdtd_dev_update_H(fdtd_dev *d, double dt, int k0, int k1, int scheme);    /* timefield3d.c:1679 */
int fdtd_dev_mur_E(fdtd_dev *d, double dt, int k0, int k1);                   /* timefield3d.c:2499 */
int fdtd_dev_mur_H(fdtd_dev *d, double dt, int k0, int k1);                   /* timefield3d.c:2275 */
int fdtd_dev_step(fdtd_dev *d, double dt, int scheme);                        /* timefield3d.c:2787 */

/* Boundary: convolutional PML (no reference counterpart; the reference has only Mur faces).
 * npml[6] = cells on faces x-lo, x-hi, y-lo, y-hi, z-lo, z-hi; grading order, target reflection
 * R0, kappa_max, alpha_max.  While enabled, fdtd_dev_step runs the PML update with PEC outer
 * walls instead of the Mur faces.  fdtd_dev_cpml_off restores the reference boundary. */
int fdtd_dev_cpml_on(fdtd_dev *d, const int npml[6], int order, double R0, double kappa_max, double alpha_max);
int fdtd_dev_cpml_off(fdtd_dev *d);
int fdtd_dev_update_E_cpml(fdtd_dev *d, double dt);
int fdtd_dev_update_H_cpml(fdtd_dev *d, double dt);
/* Kernel used by fdtd_dev_step under the PML boundary: rows = 0 selects the split E / H pair
 * (kchunk > 0 fixes their z planes per thread, 0 lets each launch choose), rows = -3 the
 * single-pass full step (E then H of one step in one z-march of 64 x 4 tiles, one launch per
 * slab class, ping-pong fields and CPML memories; kchunk = z planes per chunk, 0 automatic),
 * rows = -2 the
 * single-pass leap launches for device-resident runs (fdtd_dev_run: H of step s and E of step
 * s + 1 in one launch, each field component read and written once; kchunk = z planes per chunk,
 * 0 automatic; a lone fdtd_dev_step stays a split pair), rows = -1 the fused
 * single-kernel step in columns (one thread per (i, j) column, kchunk = 1, 2 or 4 z planes, the
 * neighbours' new E recomputed; evaluates run monitors itself), rows = 1, 2, 4 or 8 the fused
 * step in register strips with that many y rows per lane and kchunk z planes per wavefront,
 * rows = 16 or 32 the fused step in LDS tiles of 8 or 16 rows and kchunk z planes per block.
 * Default: the split pair with automatic chunking. */
int fdtd_dev_set_step_kernel(fdtd_dev *d, int rows, int kchunk);
/* Cache policy of the split updates under the PML: the loads of each cell's own old field may
 * carry the streaming (non-temporal) hint, which pays off once the fields exceed the last-level
 * cache and costs grids that fit it.  mode -1 (default): on when the six field arrays exceed
 * 240 MB; 0 off; 1 on.  A batch decides once, on its grids' total field bytes (every batched
 * launch sweeps all grids), with the first grid's mode. */
int fdtd_dev_set_streaming_loads(fdtd_dev *d, int mode);

/* Device-resident run (extension).  A run holds mode sources whose per-step amplitude and
 * phase are tabulated up front, FP64 mode-overlap monitors with a running DFT at nfreq angular
 * frequencies, and FP64 multi-frequency DFT planes.  fdtd_dev_run(n0, n) advances n steps; step
 * n writes the sources with amplitude[n]/phase[n], runs fdtd_dev_step and then accumulates the
 * monitors with twiddle e^{i omega (n + t_offset) dt}.  No host synchronisation happens inside. */
int fdtd_dev_run_clear(fdtd_dev *d);
int fdtd_dev_run_add_source(fdtd_dev *d, int cx, int cy, int cz, const double *mode, int mi, int mj, double mdisc,
                            int reverse, int adjoint, int nsteps, const double *amplitude, const double *phase);
int fdtd_dev_run_add_overlap(fdtd_dev *d, int cx, int cy, int cz, const double *mode, int mi, int mj, double mdisc,
                             int reverse, int nfreq, const double *omega);
int fdtd_dev_run_add_dft_plane(fdtd_dev *d, int k, int comp_mask, int nfreq, const double *omega);
/* DFT monitor over the box [box0,box1) x [box2,box3) x [box4,box5) (i, j, k); result layout
 * [comp][freq][k][j][i] of the box. */
int fdtd_dev_run_add_dft_box(fdtd_dev *d, const int box[6], int comp_mask, int nfreq, const double *omega);
/* Exact adjoint source of the overlap monitor (cx, cy, cz, mode, reverse): on the monitor plane
 * J = dx^2/2 (mHy', -mHx', 0) and M = e^{-i omega_dt} dx^2/2 (mEy, -mEx, 0) (mH' = -conj(mH) when
 * reverse), step n writing Re[(J, M) amplitude[n] e^{-i phase[n]}].  This is the reference's CW
 * adjoint source (timefield3d.c:3028-3068) with the overlap's scale and the one-step phase of
 * the magnetic part that make the discrete adjoint exact. */
int fdtd_dev_run_add_adjoint_source(fdtd_dev *d, int cx, int cy, int cz, const double *mode, int mi, int mj,
                                    double mdisc, int reverse, double omega_dt, int nsteps, const double *amplitude,
                                    const double *phase);
/* grad[q] = Re[(w_re + i w_im) * sum_m F_m(q) A_m(q)] over the E components m both DFT monitors
 * hold, at frequency index freq; both monitors cover the same box (q indexes it x fastest). */
int fdtd_dev_dft_gradient(fdtd_dev *df, int id_f, fdtd_dev *da, int id_a, int freq, double w_re, double w_im,
                          double *grad);
int fdtd_dev_run(fdtd_dev *d, double dt, int n0, int n, double t_offset);
/* Monitor results: overlap -> cumulative[2*nfreq] (and the last instantaneous value);
 * dft plane -> [comp][freq][J][I] complex interleaved. */
int fdtd_dev_run_overlap_result(fdtd_dev *d, int id, double *cumulative, double *last_inst);
int fdtd_dev_run_dft_result(fdtd_dev *d, int id, double *out);

/* Batched runs (extension): the grids advance in lockstep, one E and one H launch per step over
 * all of them (kernels over a device table of per-grid descriptors; steps replayed from a
 * captured hipGraph; the monitors' per-step overlaps and running DFTs of every grid are formed in
 * two launches when the call ends).  Every grid runs under the PML boundary with the split kernels and
 * holds its run configuration (at most 2 run sources and 4 overlap monitors clear of the boundary
 * ring, no DFT monitors).  fdtd_dev_batch_run(b, dt, n0, n) advances steps n0..n0+n-1 of every
 * grid's run; results are read per grid (fdtd_dev_run_overlap_result).  The launch shape (z
 * planes per thread, planes unrolled or looped, persistent or one block per work unit) is picked
 * on the first run by timing real steps with each candidate unless fdtd_dev_batch_set_shape fixed
 * it (kchunk 0 = tune); fdtd_dev_batch_info reports the shape {kchunk, unroll, persistent}, each
 * candidate's code (kchunk + 16 unroll + 32 persistent) and ms per step, and the cell count. */
typedef struct fdtd_dev_batch fdtd_dev_batch;
fdtd_dev_batch *fdtd_dev_batch_new(fdtd_dev *const *grids, int n);
void fdtd_dev_batch_free(fdtd_dev_batch *b);
int fdtd_dev_batch_set_shape(fdtd_dev_batch *b, int kchunk, int unroll, int persistent);
int fdtd_dev_batch_run(fdtd_dev_batch *b, double dt, int n0, int n);
int fdtd_dev_batch_info(fdtd_dev_batch *b, int shape[3], long long *cells, int *ntuned, int tuned_shape[16],
                        double tuned_ms[16]);
int fdtd_dev_batch_timer_start(fdtd_dev_batch *b);
int fdtd_dev_batch_timer_stop(fdtd_dev_batch *b, double *elapsed_ms);
/* Per-kernel timing of a batch: while on, steps run without the captured graph and every E / H
 * launch is bracketed by events its own dispatch records; kernel_times returns per slot (0 E,
 * 1 H) the number of timed launches and their summed durations in ms, and clears the record. */
int fdtd_dev_batch_kernel_timing(fdtd_dev_batch *b, int on);
int fdtd_dev_batch_kernel_times(fdtd_dev_batch *b, int counts[2], double ms[2]);

/* Soft sources on plane k = cz.  mode is interleaved complex double[6*mi*mj] (already
 * power-normalised by the caller).  adjoint selects the CW adjoint form (timefield3d.c:3028). */
int fdtd_dev_mode_source(fdtd_dev *d, int cx, int cy, int cz, const double *mode, int mi, int mj, double mdisc,
                         double cos_phase, double sin_phase, double amplitude, int reverse, int adjoint);
/* Overwrite plane k of the sources with a host AoS plane float[6*I*J]; cells where mask == 0
 * keep their previous value (mask may be NULL). */
int fdtd_dev_set_source_plane(fdtd_dev *d, int k, const float *plane, const unsigned char *mask);

/* Instantaneous overlap on plane cz (timefield3d.c:3176-3239, without the cumulative update). */
int fdtd_dev_mode_overlap(fdtd_dev *d, int cx, int cy, int cz, const double *mode, int mi, int mj, double mdisc,
                          int reverse, double out[2]);

/* The same instantaneous overlap without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  mode_dev is the power-normalised mode already in
 * device memory (interleaved complex double[6*mi*mj], as for fdtd_dev_mode_overlap), and out_dev is
 * 16-byte-aligned device memory that receives (re, im).  The call only enqueues on the grid's
 * stream (synchronise it before reading out_dev) and allocates nothing.  Its block partials and
 * last-block counter live in a static per-GPU scratch array: planes are limited to 8192 * 128 =
 * 1,048,576 cells (larger ones fail with -2), and calls that overlap in time on one GPU race on
 * that array. */
int fdtd_dev_mode_overlap_noalloc(fdtd_dev *d, int cx, int cy, int cz, const double *mode_dev, int mi, int mj,
                                  double mdisc, int reverse, double *out_dev);

/* Single-frequency running DFT with a host-computed twiddle; y_slice < 0 means the volume;
 * [k0, k1) bounds z (timefield3d.c:2851-2975). */
int fdtd_dev_dft(fdtd_dev *d, double tw_re, double tw_im, int y_slice, int k0, int k1);

/* Taper edge histories, keyed by an opaque host pointer (adjoint3d.c:259-373). */
int fdtd_dev_taper_store(fdtd_dev *d, const void *key, int which, int seg_num, int start, int end, const int *left,
                         const int *right, int bottom, int top, int duration, int timestep);
int fdtd_dev_taper_ep_gradient(const void *key, int height, int length, int duration, double *out);
void fdtd_dev_taper_release(const void *key);

/* 3-D block adjoint (adjoint3d.c:48-105), keyed by an opaque host pointer. */
int fdtd_dev_block_store(fdtd_dev *d, const void *key, const int box[6], int duration, int timestep);
int fdtd_dev_block_update_gradient(fdtd_dev *d, const void *key, const int box[6], int duration, int timestep,
                                   double *gradient_host);
int fdtd_dev_block_reset(const void *key, int which);
void fdtd_dev_block_release(const void *key);

/* Per-kernel timing of the Yee updates: while enabled, every `every`-th launch of each update
 * kernel is bracketed by HIP events on the grid's stream; kernel_times returns, for the split E
 * kernel, the split H kernel and the fused step kernel (slots 0, 1, 2), the number of sampled
 * launches and their summed durations in ms, and resets the sample. */
int fdtd_dev_kernel_timing(fdtd_dev *d, int every);
int fdtd_dev_kernel_times(fdtd_dev *d, int counts[3], double ms[3]);

/* Non-finite field detection (no counterpart in the reference, whose loops run on through NaN
 * and Inf).  count_nonfinite scans the six field arrays once (synchronous) and returns the number
 * of NaN / Inf values.  With set_finite_check(d, every > 0), fdtd_dev_run scans the fields after
 * every step whose index s has (s + 1) % every == 0 and, after finishing the call, returns -4
 * with the count and the first step that showed them in fdtd_dev_last_error() (0 = off). */
int fdtd_dev_count_nonfinite(fdtd_dev *d, long long *count);
int fdtd_dev_set_finite_check(fdtd_dev *d, int every);

/* Halo planes of z-slab shards (no counterpart in the reference, which runs one grid): copy the
 * components in comp_mask (bit m = component m: Ex, Ey, Ez, Hx, Hy, Hz) of plane ks of src into
 * plane kd of dst (same I, J), ordered after src's queued work and before either grid's later
 * work.  A shard keeps its neighbours' planes in its boundary planes, which no kernel updates. */
int fdtd_dev_copy_planes(fdtd_dev *dst, int kd, fdtd_dev *src, int ks, unsigned comp_mask);

/* Timing of the work enqueued on the grid's stream (HIP events on that stream). */
int fdtd_dev_timer_start(fdtd_dev *d);
int fdtd_dev_timer_stop(fdtd_dev *d, double *elapsed_ms);
int fdtd_dev_synchronize(fdtd_dev *d);

/* Device memory and streams released by fdtd_dev_destroy / fdtd_dev_batch_free / fdtd_mlp_free
 * are cached for the next grid (a design loop creates and destroys grids by the hundred); this
 * returns every cached block to the HIP runtime.  No counterpart in the reference (its grids are
 * host memory). */
int fdtd_dev_release_cached_memory(void);

#ifdef __cplusplus
}
#endif
#endif
