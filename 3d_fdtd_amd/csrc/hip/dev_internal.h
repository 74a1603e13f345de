// Internal declarations shared by the HIP translation units of the device layer.
// Not part of the C-ABI: the public entry points are declared in include/fdtd3d_dev.h.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "fdtd3d_dev.h"

namespace fdtd {

// Device memory and streams of the device layer come from a cache (dev_alloc.hip): a released
// block is reused after one device synchronisation, so release it only once work that uses it
// has been enqueued (never while the host still needs its content); release a stream only after
// synchronising it.
hipError_t dev_malloc_bytes(void **p, size_t bytes);
template <class T>
inline hipError_t dev_malloc(T **p, size_t bytes) {
    return dev_malloc_bytes(reinterpret_cast<void **>(p), bytes);
}
hipError_t dev_free(void *p);
// Rewrite grid d's material-id volume with `bytes`-byte ids (no-op when already that wide).
int widen_material_ids(fdtd_dev *d, int bytes);
hipError_t dev_stream_acquire(hipStream_t *s);
hipError_t dev_stream_release(hipStream_t s);

// Component order inside a cell: Ex, Ey, Ez, Hx, Hy, Hz (timefield3d.h:19-25).
enum { EX = 0, EY, EZ, HX, HY, HZ };

constexpr int kMaxSrcPlanes = 8;
// Threads per block of the stencil, monitor and reduction kernels.  Two-wavefront blocks measured
// fastest for the split E / H updates on the strip waveguide (H update 22.8 us at 128 threads,
// 23.1 at 64, 23.0 at 192, 23.2 at 256, 25.2 at 512; same run, same grid).
constexpr int kBlock = 128;
// Grids with fewer cells than this run the split and batched kernels with 32-bit cell offsets
// (every per-grid array is then below 4 GiB).
constexpr long long kSmallGridCells = 1LL << 30;
// Streaming (non-temporal) loads of the own-cell fields in the split updates pay off once the
// six field arrays no longer fit the 256 MB last-level cache and cost grids that fit it (200 x
// 100 cross-section, E + H per step: 3 M cells 0.87x, 6 M 0.83x, 9 M 0.99x, 12.4 M 1.14x, 18 M
// 1.06x).  mode: -1 automatic, 0 off, 1 on.
constexpr long long kStreamFieldBytes = 240LL << 20;
inline bool streaming_loads(long long cells, int mode) {
    return mode > 0 || (mode < 0 && 24LL * cells > kStreamFieldBytes);
}

// Per-material coefficient record (floats), one for each conductivity scheme:
//   [0..2] caE, [3..5] cbE, [6..8] cjE, [9..11] caH, [12..14] cbH, [15..17] cjH,
//   [18..23] Mur s for E on faces x,y,z (two tangential comps each, order b=(a+1)%3, c=(a+2)%3),
//   [24..29] Mur s for H, [30..31] pad.
constexpr int kCoefStride = 32;
// Behind the two scheme tables the same allocation holds a packed copy of the semi-implicit
// update coefficients, kCoefPacked floats per material: caEx caEy caEz cbEx | cbEy cbEz caHx caHy |
// caHz cbHx cbHy cbHz (three 16-byte loads per cell for the single-pass step).
constexpr int kCoefPacked = 12;
// Then one float per material, its cbE (meaningful when the table is isotropic, see coef_iso).
inline size_t coef_table_floats(int nmat) { return (size_t)nmat * (2 * kCoefStride + kCoefPacked + 1); }
inline size_t coef_packed_offset(int nmat) { return (size_t)2 * nmat * kCoefStride; }
inline size_t coef_iso_offset(int nmat) { return (size_t)nmat * (2 * kCoefStride + kCoefPacked); }
// Raw per-material properties kept for sources and reports: ep[3], mu[3], sig[6] (+4 pad).
constexpr int kPropStride = 16;

struct SrcPlaneArg {
    int k;
    const float *J;  // 6 SoA planes of I*J floats (Jx,Jy,Jz,Mx,My,Mz)
};

// A tabulated mode source evaluated inside the Yee kernels during a device-resident run:
// inside the mode window J_m = c1 PR_m + c2 PI_m (PR/PI: 6 x P planes), outside it the stored
// source plane applies.
constexpr int kMaxRunSrc = 2;
struct RunSrcArg {
    int k;
    const float *PR, *PI;
    const unsigned char *win;
    float c1, c2;
    // when tab is set (graph-captured batch steps), (c1, c2) = tab[step] for 0 <= step < ntab,
    // else 0, with step read from the launch's device step counter
    const float2 *tab;
    int ntab;
};

// An overlap monitor evaluated inside the H update of a device-resident run: each block of the
// monitor plane stores its FP64 partial of Hx mEy - Hy mEx + Ey mHx - Ex mHy (W holds Re/Im of
// mEx, mEy, mHx, mHy as 8 x P floats) into part[block] for the step; the running DFT over the
// steps of a run call is formed afterwards from those partials in a fixed order.
constexpr int kMaxRunMon = 4;
struct RunMonArg {
    int k;
    const float *W;
    double2 *part;     // this step's [plane blocks] partials
    long long stride;  // with a device step counter: part + (step - step0) * stride is the step's row
};

struct SrcPlane {
    int k;
    float *J;          // device; component m of cell p at J[m * stride + p]
    long long stride;  // P for a stand-alone plane, N inside the dense source volume
};

struct ModeCache {
    const double *host_key = nullptr;
    int mi = 0, mj = 0;
    double mdisc = 0;
    std::vector<double> shadow;  // last uploaded content
    double *dev = nullptr;       // 12*mi*mj doubles (interleaved complex)
    unsigned long long last_use = 0;
    unsigned long long version = 0;  // stamp of the last upload (unique per grid)
};

// One plane-overlap request of the drop-in surface (timefield3d_mode_overlap's arguments).
struct OvlReq {
    int cx, cy, cz, mi, mj, reverse;
    double mdisc;
    const double *host_key;
};
// Overlap requests evaluated by one launch batch (one synchronisation).
constexpr int kOvlSlots = 8;

struct TaperHist {
    const void *key;
    int H, L, T;
    float *hist[2];              // forward, adjoint: [T][L][H][2][6]
    int *edge_i;                 // [2][L] edge pixel per z row (right, left)
    std::vector<int> edge_host;  // cached copy of edge_i
};

// Convolutional PML (an extension; the reference has only Mur faces).  Along axis a the slab
// index of coordinate n is n for n < nlo and nlo + n - (N-1-nhi) for n >= N-1-nhi (W = nlo +
// nhi + 1 entries, or nlo when nhi == 0); psi arrays replace coordinate a by that index.
struct Cpml {
    bool on = false;
    int nlo[3] = {0, 0, 0}, nhi[3] = {0, 0, 0}, W[3] = {0, 0, 0};
    int order = 3;
    double R0 = 1e-8, kappa_max = 1.0, alpha_max = 0.0;
    float *tab = nullptr;  // per axis: bE, aE, kE, bH, aH, kH, each N_axis floats
    long long tab_off[3] = {0, 0, 0};
    double tab_dt = -1.0;
    float *psi_base = nullptr;
    float *psi_alt = nullptr;  // second psi buffer for the fused (ping-pong) step
    float *psi[2][3][2] = {};  // [E/H][slab axis][q]: component (axis + 1 + q) % 3
    long long psi_n[3] = {0, 0, 0};
};

// Sampled per-kernel timing of the E / H updates.  A sampled launch goes through
// hipExtLaunchKernelGGL with a start/stop event pair, which the dispatch itself records, so the
// interval is the kernel's own execution (no marker packets around it).
struct KernelTiming {
    int every = 0;
    long long launches[3] = {0, 0, 0};
    std::vector<hipEvent_t> ev[3];  // start/stop pairs: split E, split H, fused step
};

// Events for the next launch of kernel slot `which`: a fresh pair on a sampled launch, else null.
void timing_events(fdtd_dev *d, int which, hipEvent_t *start, hipEvent_t *stop);

struct BlockHist {
    const void *key;
    int box[6];
    int T;
    float *fwd;     // [T][nz][ny][nx][3]
    double *grad;   // [nz][ny][nx]
};

}  // namespace fdtd

struct fdtd_dev {
    int I, J, K;
    long long P;  // cells per z plane
    long long N;  // cells
    long long Nal;  // stride between component arrays (N rounded up to 64)
    double disc;
    hipStream_t stream;

    float *field_base;  // one allocation, 6 component arrays of N floats
    float *F[6];
    // second field buffer for the fused (ping-pong) step; alt_valid = its wall cells match
    float *alt_base;
    bool alt_valid;
    int fused_rows;     // rows per lane of the fused step (0 = split E/H kernels)
    int fused_kchunk;   // z planes per wavefront of the fused step
    int split_kchunk;   // z planes per thread of the split E/H kernels (0 = chosen per launch)
    // single-pass leap launches of device-resident runs (fused_rows == -2): z planes per chunk
    // (0 = chosen per grid)
    int leap_kc;
    // single-pass full steps (fused_rows == -3): x-tile width (0: the default, 64) and z planes per
    // chunk (0 = chosen per grid)
    int onepass_tw, onepass_kc;

    // materials
    int nmat;
    int mat_bytes;        // 1, 2 or 4 (4: split kernels only)
    long long mat_uploads = 0, mat_paints = 0;  // full material uploads, device repaints
    void *mat;            // per-cell id
    // an extruded geometry's ids (set_materials_extruded): ids[k][i] in rows [ext_j0, ext_j1),
    // ext_bg elsewhere; kept while the id volume equals it (the batched kernels then read this
    // K x I table instead of the volume), null once a full upload or a repaint changes the ids
    uint16_t *ext_ids = nullptr;
    int ext_j0 = 0, ext_j1 = 0, ext_bg = 0;
    std::vector<float> props;  // nmat * kPropStride
    float *d_props;
    float *d_coef;        // [2 schemes][nmat][kCoefStride], then the packed and isotropic tables
    double coef_dt;       // dt the coefficient table was built for (NaN = invalid)
    // every material of the semi-implicit table has caE = caH = 1 on all components, one cbE for
    // its three components and the grid one cbH (isotropic, lossless, mu uniform): the update then
    // needs only cbE per material (coef_iso_offset) and the scalar cbh
    bool coef_iso;
    float coef_cbh;
    bool have_mats;

    // Mur shells: 6 faces x 2 layers x 4 comps x face area
    float *shell;
    long long shell_off[6];

    // sources: up to kMaxSrcPlanes z planes; beyond that a dense 6 x N volume (dense_src)
    std::vector<fdtd::SrcPlane> src;
    float *dense_src;
    fdtd::SrcPlane dense_view;

    // reference single-frequency DFT accumulators (6 SoA arrays of N double2), lazily allocated
    double2 *fourier;

    // overlap reduction scratch
    double2 *partials;
    double2 *red_out;
    int partial_cap;
    unsigned *ovl_count;  // last-block counters of k_overlap, one per slot (zero between launches)
    double2 *ovl_host;    // pinned, coherent host memory the overlap results land in (one per slot)
    double2 *ovl_dev;     // its device address
    // Overlap batching: the fields change only through calls that advance field_epoch.  The
    // requests seen since the last change (ovl_seen) become the pattern (ovl_prev) that the first
    // request after the next change evaluates along with itself; their results stay valid while
    // field_epoch == batch_epoch and the mode content is unchanged (same upload version).
    unsigned long long field_epoch, mode_uploads;
    fdtd::OvlReq ovl_seen[fdtd::kOvlSlots], ovl_prev[fdtd::kOvlSlots], ovl_req[fdtd::kOvlSlots];
    unsigned long long ovl_ver[fdtd::kOvlSlots];
    int n_seen, n_prev, n_batch;
    unsigned long long seen_epoch, batch_epoch;

    std::vector<fdtd::ModeCache> modes;
    unsigned long long use_clock;

    fdtd::Cpml cpml;
    void *run;  // device-resident run configuration (dev_run.hip), or null
    int n_run_src;                              // run sources evaluated by the next step
    fdtd::RunSrcArg run_src[fdtd::kMaxRunSrc];
    int n_run_mon;                              // overlap monitors folded into the next H update
    fdtd::RunMonArg run_mon[fdtd::kMaxRunMon];
    fdtd::KernelTiming ktime;

    // non-finite field watch of device-resident runs (dev_health.hip): every `finite_every` steps
    // a scan adds the count of NaN / Inf field values and the first step that showed any to
    // finite_buf (unsigned long long count, int first step)
    int finite_every;
    void *finite_buf;
    int stream_mode;  // streaming own-cell loads in the split updates: -1 automatic, 0 off, 1 on

    hipEvent_t ev0, ev1;
    hipEvent_t xev;  // ordering event of halo-plane copies between grids (no timing)
};

namespace fdtd {

void set_error(const std::string &msg);
int fail(const char *where, hipError_t e);
#define FDTD_TRY(expr)                                              \
    do {                                                            \
        hipError_t _e = (expr);                                     \
        if (_e != hipSuccess) return ::fdtd::fail(#expr, _e);       \
    } while (0)

int ensure_coefficients(fdtd_dev *d, double dt);
int cpml_zero(fdtd_dev *d);
int cpml_tables(fdtd_dev *d, double dt);
int fused_step(fdtd_dev *d, double dt, int ty_rows, int kc_len);
bool fused_supported(const fdtd_dev *d);
// True when fdtd_dev_step runs the split E/H kernels (the path that can fold run monitors).
bool step_uses_split(const fdtd_dev *d);
int fused_sync_alt(fdtd_dev *d);
int fused_prepare(fdtd_dev *d);
void fused_swap(fdtd_dev *d);
// Fused one-kernel step in columns (dev_colstep.hip): kc z planes per thread (1, 2 or 4).
int col_step(fdtd_dev *d, double dt, int kc);
// Non-finite watch of runs (dev_health.hip): reset before a run call, scan after step `step`,
// and the verdict after the call (0, or -4 with the count and first step in the message).
int finite_watch_reset(fdtd_dev *d);
int finite_watch_scan(fdtd_dev *d, int step);
int finite_watch_verdict(fdtd_dev *d);
bool col_supported(const fdtd_dev *d);
// The step kernel evaluates run overlap monitors itself (split H update or column step).
bool step_folds_monitors(const fdtd_dev *d);
// Monitor partial slots per step that a folding step kernel writes (one per block of a plane).
int step_monitor_slots(const fdtd_dev *d);
// True when fdtd_dev_step runs the single-pass full step (dev_onepass.hip).
bool step_uses_onepass(const fdtd_dev *d);
// The device source plane for z = k (created zeroed on first use), or null with *err set.
SrcPlane *plane_for(fdtd_dev *d, int k, int *err);
void run_release(fdtd_dev *d);
const double *mode_on_device(fdtd_dev *d, const double *mode, int mi, int mj, double mdisc, int *err);
// The cached device copy of a mode as last uploaded (no upload, no content check), or null.
const ModeCache *mode_cached(const fdtd_dev *d, const double *mode, int mi, int mj, double mdisc);
// Every call that may change the field values advances the grid's field epoch (overlap batching).
inline void fields_changed(fdtd_dev *d) { d->field_epoch++; }

// Blocks over (plane chunks x z chunks) remapped so that each XCD walks a contiguous range
// (workgroups are dispatched round-robin over the 8 XCDs of an MI355X).
__device__ __forceinline__ int xcd_logical_block(int bid, int nb) {
    int x = bid & 7, s = bid >> 3, q = nb >> 3, r = nb & 7;
    return x * q + (x < r ? x : r) + s;
}
__device__ __forceinline__ long long xcd_logical_block_ll(long long bid, long long nb) {
    long long x = bid & 7, s = bid >> 3, q = nb >> 3, r = nb & 7;
    return x * q + (x < r ? x : r) + s;
}

}  // namespace fdtd
