// Batched multi-geometry runs: many grids of uneven sizes (e.g. taper candidates of different
// lengths) advance in lockstep, one E launch and one H launch per step covering every grid.
//
// Each grid keeps its own device state (fdtd_dev) and run configuration (tabulated sources,
// overlap monitors).  A device table holds one descriptor per grid (the split-kernel arguments
// plus the grid's first virtual block); the launches cover the concatenated virtual blocks of all
// grids, one block per virtual block or a persistent grid striding over them, remapped so that
// each XCD walks a contiguous range (consecutive z chunks of one grid share that XCD's L2).
// Per-step values (source amplitudes, monitor partial slots) are read through a device step
// counter, so a segment of steps (E, H, counter tick) is captured once in a hipGraph and
// replayed.  The launch shape (z-chunk depth, unrolled or looped chunk, persistent or not) is
// picked by a small autotuner that times real steps of the run with each candidate.  When every
// grid is an extruded geometry the kernels read material ids from the grids' K x I tables.
#include <cstring>
#include <vector>

#include "dev_stencil.h"

namespace fdtd {
namespace {

struct BatchGeom {
    YeeArgs a;
    PmlArgs pm;
    long long vb0;  // first virtual block of this grid
};

// The grid holding virtual block vb (descriptors sorted by vb0).
__device__ __forceinline__ int batch_grid(const BatchGeom *__restrict__ geo, int G, long long vb) {
    int lo = 0, hi = G - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (geo[mid].vb0 <= vb) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// NT: streaming own-cell loads (the batch's fields exceed the last-level cache).  LOOP: a
// persistent grid-stride launch; else one block per virtual block, with no state carried from
// one virtual block to the next (fewer registers).
template <typename MatT, int KC, bool NT, bool LOOP>
__global__ __launch_bounds__(kBlock) void k_batch_E(const BatchGeom *__restrict__ geo, int G, long long total) {
    auto block = [&](long long vb) {
        const long long lb = xcd_logical_block_ll(vb, total);
        const BatchGeom &B = geo[batch_grid(geo, G, lb)];
        update_E_block<MatT, true, KC, uint32_t, NT>(B.a, B.pm, (int)(lb - B.vb0));
    };
    if constexpr (LOOP) {
        for (long long vb = blockIdx.x; vb < total; vb += gridDim.x) block(vb);
    } else if (blockIdx.x < total) {
        block(blockIdx.x);
    }
}

template <typename MatT, bool MON, int KC, bool NT, bool LOOP>
__global__ __launch_bounds__(kBlock) void k_batch_H(const BatchGeom *__restrict__ geo, int G, long long total) {
    auto block = [&](long long vb) {
        const long long lb = xcd_logical_block_ll(vb, total);
        const BatchGeom &B = geo[batch_grid(geo, G, lb)];
        update_H_block<MatT, true, MON, KC, uint32_t, NT>(B.a, B.pm, (int)(lb - B.vb0));
    };
    if constexpr (LOOP) {
        for (long long vb = blockIdx.x; vb < total; vb += gridDim.x) {
            block(vb);
            if (MON) __syncthreads();  // the next virtual block reuses the reduction slots
        }
    } else if (blockIdx.x < total) {
        block(blockIdx.x);
    }
}

__global__ void k_step_tick(int *step) {
    if (threadIdx.x == 0) *step += 1;
}

__global__ void k_step_set(int *step, int v) {
    if (threadIdx.x == 0) *step = v;
}

}  // namespace
}  // namespace fdtd

using namespace fdtd;

struct fdtd_dev_batch {
    std::vector<fdtd_dev *> grids;
    hipStream_t stream = nullptr;
    int *step = nullptr;        // device step counter
    BatchStage *stage = nullptr;  // pinned staging and device tables of a call's finish
    BatchGeom *geo = nullptr;   // device descriptor table
    std::vector<BatchGeom> host;
    long long total = 0;        // virtual blocks of the current chunk depth
    // launch shape: z planes per thread, unrolled (all planes loaded first) or a runtime loop,
    // persistent grid (2048 blocks) or one block per virtual block; kc = 0: tune on the next run
    int kc = 0;
    bool unroll = true, persistent = true;
    bool u8 = true, mon = false;
    bool ext = false;  // every grid extruded: ids from the K x I tables, not the id volumes
    bool nt = false;  // streaming own-cell loads: the batch's fields exceed the cache
    unsigned grid_blocks = 0;
    // captured segment of kSeg steps
    hipGraphExec_t exec = nullptr;
    int exec_kc = -1;
    bool exec_unroll = false, exec_persistent = false, exec_mon = false, exec_ext = false;
    long long exec_total = -1;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // per-kernel timing: while on, steps run without the captured graph and every E / H launch
    // carries a start/stop event pair recorded by its own dispatch
    bool ktime = false;
    std::vector<hipEvent_t> kev[2];
    static constexpr int kMaxTuned = 16;
    double tuned_ms[kMaxTuned] = {};
    int tuned_code[kMaxTuned] = {};  // kc + 16 * unroll + 32 * persistent
    int ntuned = 0;
};

namespace {

constexpr int kSeg = 32;           // steps per captured graph segment
constexpr int kBatchChunk = 4096;  // steps per prepare/finish cycle (bounds the monitor buffers)
struct Shape {
    int kc;
    bool unroll, persistent;
};
const Shape kTune[] = {{2, true, true},  {4, true, true},  {2, false, true},  {4, false, true},  {8, false, true},
                       {2, true, false}, {4, true, false}, {2, false, false}, {4, false, false}, {8, false, false},
                       {3, true, false}};

long long batch_layout(fdtd_dev_batch *b, int kc) {
    long long vb = 0;
    for (auto &g : b->host) {
        g.a.kchunk = kc;
        const int nz = g.a.k1 - g.a.k0;
        const long long zchunks = (nz + kc - 1) / kc;
        g.vb0 = vb;
        vb += (long long)g.a.plane_blocks * zchunks;
    }
    return vb;
}

int upload_layout(fdtd_dev_batch *b, int kc, bool persistent) {
    b->total = batch_layout(b, kc);
    FDTD_TRY(hipMemcpyAsync(b->geo, b->host.data(), b->host.size() * sizeof(BatchGeom), hipMemcpyHostToDevice,
                            b->stream));
    FDTD_TRY(hipStreamSynchronize(b->stream));  // the host table is rewritten by the next layout
    // persistent: 2048 threads per CU on 256 CUs; else one block per virtual block; either way
    // a multiple of 8 (XCDs) so that a grid-stride step keeps the XCD
    const long long cap = 256LL * 2048 / kBlock;
    long long g = persistent && b->total > cap ? cap : b->total;
    g = (g + 7) / 8 * 8;
    b->grid_blocks = (unsigned)(g < 8 ? 8 : g);
    return 0;
}

// A start/stop event pair for the next timed launch of kernel slot `which` (0 E, 1 H), or nulls.
void batch_events(fdtd_dev_batch *b, int which, hipEvent_t *t0, hipEvent_t *t1) {
    *t0 = *t1 = nullptr;
    if (!b->ktime) return;
    hipEvent_t x = nullptr, y = nullptr;
    if (hipEventCreate(&x) != hipSuccess) return;
    if (hipEventCreate(&y) != hipSuccess) {
        (void)hipEventDestroy(x);
        return;
    }
    b->kev[which].push_back(x);
    b->kev[which].push_back(y);
    *t0 = x;
    *t1 = y;
}

template <typename MatT, int KC, bool NT, bool LOOP>
void launch_step_typed(fdtd_dev_batch *b) {
    const int G = (int)b->host.size();
    const dim3 grid(b->grid_blocks), block(kBlock);
    const BatchGeom *geo = b->geo;
    if (b->ktime) {
        hipEvent_t e0, e1, h0, h1;
        batch_events(b, 0, &e0, &e1);
        batch_events(b, 1, &h0, &h1);
        hipExtLaunchKernelGGL((k_batch_E<MatT, KC, NT, LOOP>), grid, block, 0, b->stream, e0, e1, 0, geo, G, b->total);
        if (b->mon)
            hipExtLaunchKernelGGL((k_batch_H<MatT, true, KC, NT, LOOP>), grid, block, 0, b->stream, h0, h1, 0, geo, G,
                                  b->total);
        else
            hipExtLaunchKernelGGL((k_batch_H<MatT, false, KC, NT, LOOP>), grid, block, 0, b->stream, h0, h1, 0, geo, G,
                                  b->total);
    } else {
        hipLaunchKernelGGL((k_batch_E<MatT, KC, NT, LOOP>), grid, block, 0, b->stream, geo, G, b->total);
        if (b->mon)
            hipLaunchKernelGGL((k_batch_H<MatT, true, KC, NT, LOOP>), grid, block, 0, b->stream, geo, G, b->total);
        else
            hipLaunchKernelGGL((k_batch_H<MatT, false, KC, NT, LOOP>), grid, block, 0, b->stream, geo, G, b->total);
    }
}

template <typename MatT, int KC, bool NT>
void launch_step_loop(fdtd_dev_batch *b) {
    // the grid covers every virtual block exactly when it is not capped at the persistent size
    if (b->grid_blocks < b->total) launch_step_typed<MatT, KC, NT, true>(b);
    else launch_step_typed<MatT, KC, NT, false>(b);
}

template <int KC>
void launch_step_kc(fdtd_dev_batch *b) {
    if (b->ext && b->nt) launch_step_loop<ExtIds, KC, true>(b);
    else if (b->ext) launch_step_loop<ExtIds, KC, false>(b);
    else if (b->u8 && b->nt) launch_step_loop<uint8_t, KC, true>(b);
    else if (b->u8) launch_step_loop<uint8_t, KC, false>(b);
    else if (b->nt) launch_step_loop<uint16_t, KC, true>(b);
    else launch_step_loop<uint16_t, KC, false>(b);
    hipLaunchKernelGGL(k_step_tick, dim3(1), dim3(64), 0, b->stream, b->step);
}

int launch_step(fdtd_dev_batch *b, int kc, bool unroll) {
    switch (unroll ? kc : 0) {
        case 2: launch_step_kc<2>(b); break;
        case 3: launch_step_kc<3>(b); break;
        case 4: launch_step_kc<4>(b); break;
        default: launch_step_kc<0>(b); break;
    }
    FDTD_TRY(hipGetLastError());
    return 0;
}

// The captured segment for the current layout (re-captured when the kernel choice changes).
int segment_graph(fdtd_dev_batch *b) {
    if (b->exec && b->exec_kc == b->kc && b->exec_unroll == b->unroll && b->exec_persistent == b->persistent &&
        b->exec_mon == b->mon && b->exec_ext == b->ext && b->exec_total == b->total)
        return 0;
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    hipGraph_t graph = nullptr;
    FDTD_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int s = 0; s < kSeg && rc == 0; s++) rc = launch_step(b, b->kc, b->unroll);
    hipError_t e = hipStreamEndCapture(b->stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail("hipStreamEndCapture", e);
    e = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail("hipGraphInstantiate", e);
    b->exec_kc = b->kc;
    b->exec_unroll = b->unroll;
    b->exec_persistent = b->persistent;
    b->exec_mon = b->mon;
    b->exec_ext = b->ext;
    b->exec_total = b->total;
    return 0;
}

int set_step(fdtd_dev_batch *b, int n0) {
    hipLaunchKernelGGL(k_step_set, dim3(1), dim3(64), 0, b->stream, b->step, n0);
    FDTD_TRY(hipGetLastError());
    return 0;
}

// Steps n0..n0+n-1 of every grid (n <= kBatchChunk); the first call of an untuned batch spends its
// first steps timing each chunk-depth candidate (real steps: results do not depend on the depth).
int batch_chunk(fdtd_dev_batch *b, double dt, int n0, int n) {
    for (auto *d : b->grids) fields_changed(d);
    for (auto *d : b->grids) FDTD_TRY(hipStreamSynchronize(d->stream));
    FDTD_TRY(hipStreamSynchronize(b->stream));  // nothing of an earlier call (failed or not) in flight
    b->mon = false;
    b->ext = true;
    for (auto *d : b->grids) b->ext = b->ext && d->ext_ids != nullptr;
    for (size_t g = 0; g < b->grids.size(); g++) {
        int rc = run_batch_prepare(b->grids[g], dt, n0, n, b->step, &b->host[g].a, &b->host[g].pm);
        if (rc) return rc;
        b->mon = b->mon || b->host[g].a.nmon > 0;
    }
    int rc = set_step(b, n0);
    if (rc) return rc;
    int done = 0;
    if (b->kc == 0) {
        const int per = 4;  // timed steps per candidate (after one untimed step)
        Shape best = kTune[0];
        double best_ms = 1e30;
        b->ntuned = 0;
        for (const Shape &c : kTune) {
            if (done + per + 1 > n || b->ntuned >= fdtd_dev_batch::kMaxTuned) break;
            if ((rc = upload_layout(b, c.kc, c.persistent))) return rc;
            if ((rc = launch_step(b, c.kc, c.unroll))) return rc;
            FDTD_TRY(hipEventRecord(b->ev0, b->stream));
            for (int s = 0; s < per; s++)
                if ((rc = launch_step(b, c.kc, c.unroll))) return rc;
            FDTD_TRY(hipEventRecord(b->ev1, b->stream));
            FDTD_TRY(hipEventSynchronize(b->ev1));
            float ms = 0;
            FDTD_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
            b->tuned_code[b->ntuned] = c.kc + 16 * c.unroll + 32 * c.persistent;
            b->tuned_ms[b->ntuned++] = ms / per;
            if (ms < best_ms) {
                best_ms = ms;
                best = c;
            }
            done += per + 1;
        }
        b->kc = best.kc;
        b->unroll = best.unroll;
        b->persistent = best.persistent;
    }
    if ((rc = upload_layout(b, b->kc, b->persistent))) return rc;
    const int left = n - done;
    const int segs = b->ktime ? 0 : left / kSeg;
    if (segs > 0) {
        if ((rc = segment_graph(b))) return rc;
        for (int s = 0; s < segs; s++) FDTD_TRY(hipGraphLaunch(b->exec, b->stream));
    }
    for (int s = segs * kSeg; s < left; s++)
        if ((rc = launch_step(b, b->kc, b->unroll))) return rc;
    if ((rc = run_batch_finish(b->grids.data(), (int)b->grids.size(), dt, n0, n, b->stream, b->stage))) return rc;
    FDTD_TRY(hipStreamSynchronize(b->stream));
    return 0;
}

}  // namespace

extern "C" {

fdtd_dev_batch *fdtd_dev_batch_new(fdtd_dev *const *grids, int n) {
    if (n <= 0) {
        set_error("batch: no grids");
        return nullptr;
    }
    int mb = 1;
    for (int g = 0; g < n; g++) {
        if (!grids[g] || grids[g]->mat_bytes > 2) {
            set_error("batch: null grid, or more than 65536 materials in a grid");
            return nullptr;
        }
        mb = grids[g]->mat_bytes > mb ? grids[g]->mat_bytes : mb;
    }
    // one id width for the whole batch: grids with 1-byte ids join with 2-byte ids
    for (int g = 0; g < n; g++)
        if (widen_material_ids(grids[g], mb)) return nullptr;
    for (int g = 0; g < n; g++) {
        if (grids[g]->N >= kSmallGridCells) {
            set_error("batch: every grid must have fewer than 2^30 cells");
            return nullptr;
        }
    }
    fdtd_dev_batch *b = new fdtd_dev_batch();
    b->grids.assign(grids, grids + n);
    b->host.resize((size_t)n);
    std::memset((void *)b->host.data(), 0, b->host.size() * sizeof(BatchGeom));
    b->u8 = grids[0]->mat_bytes == 1;
    // every launch sweeps all grids, so nothing one launch reads is still cached at the next
    // unless the whole batch fits: decide on the batch's total (64 tapers, 608 M cells, chosen
    // launch shape: 9.61-9.65 ms per step streaming against 9.69-9.72 plain)
    long long cells = 0;
    for (int g = 0; g < n; g++) cells += grids[g]->N;
    b->nt = streaming_loads(cells, grids[0]->stream_mode);
    b->stage = batch_stage_new();
    hipError_t e = dev_stream_acquire(&b->stream);
    if (e == hipSuccess) e = dev_malloc(&b->step, sizeof(int));
    if (e == hipSuccess) e = dev_malloc(&b->geo, (size_t)n * sizeof(BatchGeom));
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e != hipSuccess) {
        fail("batch create", e);
        fdtd_dev_batch_free(b);
        return nullptr;
    }
    return b;
}

void fdtd_dev_batch_free(fdtd_dev_batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    for (auto &v : b->kev)
        for (hipEvent_t e : v) (void)hipEventDestroy(e);
    (void)dev_free(b->step);
    (void)dev_free(b->geo);
    batch_stage_free(b->stage);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)dev_stream_release(b->stream);
    delete b;
}

int fdtd_dev_batch_set_shape(fdtd_dev_batch *b, int kchunk, int unroll, int persistent) {
    if (kchunk < 0 || (unroll && kchunk > 4)) {
        set_error("batch: z chunk must be >= 0 (0 = tune on the next run), unrolled chunks at most 4");
        return -2;
    }
    b->kc = kchunk;
    b->unroll = unroll != 0;
    b->persistent = persistent != 0;
    return 0;
}

int fdtd_dev_batch_run(fdtd_dev_batch *b, double dt, int n0, int n) {
    for (int s = 0; s < n; s += kBatchChunk) {
        const int rc = batch_chunk(b, dt, n0 + s, n - s < kBatchChunk ? n - s : kBatchChunk);
        if (rc) return rc;
    }
    return 0;
}

int fdtd_dev_batch_info(fdtd_dev_batch *b, int shape[3], long long *cells, int *ntuned, int tuned_shape[16],
                        double tuned_ms[16]) {
    long long c = 0;
    for (auto *d : b->grids) c += d->N;
    if (shape) {
        shape[0] = b->kc;
        shape[1] = b->unroll;
        shape[2] = b->persistent;
    }
    if (cells) *cells = c;
    if (ntuned) *ntuned = b->ntuned;
    for (int q = 0; q < b->ntuned && q < 16; q++) {
        if (tuned_shape) tuned_shape[q] = b->tuned_code[q];
        if (tuned_ms) tuned_ms[q] = b->tuned_ms[q];
    }
    return 0;
}

int fdtd_dev_batch_timer_start(fdtd_dev_batch *b) {
    FDTD_TRY(hipEventRecord(b->ev0, b->stream));
    return 0;
}

int fdtd_dev_batch_timer_stop(fdtd_dev_batch *b, double *elapsed_ms) {
    FDTD_TRY(hipEventRecord(b->ev1, b->stream));
    FDTD_TRY(hipEventSynchronize(b->ev1));
    float ms = 0;
    FDTD_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    *elapsed_ms = ms;
    return 0;
}

int fdtd_dev_batch_kernel_timing(fdtd_dev_batch *b, int on) {
    b->ktime = on != 0;
    return 0;
}

int fdtd_dev_batch_kernel_times(fdtd_dev_batch *b, int counts[2], double ms[2]) {
    FDTD_TRY(hipStreamSynchronize(b->stream));
    int rc = 0;
    for (int w = 0; w < 2; w++) {
        counts[w] = 0;
        ms[w] = 0.0;
        auto &v = b->kev[w];
        for (size_t q = 0; q + 1 < v.size(); q += 2) {
            float t = 0;
            if (hipEventElapsedTime(&t, v[q], v[q + 1]) == hipSuccess) {
                counts[w]++;
                ms[w] += t;
            } else {
                rc = fail("batch kernel times", hipErrorInvalidValue);
            }
        }
        for (hipEvent_t e : v) (void)hipEventDestroy(e);
        v.clear();
    }
    return rc;
}

}  // extern "C"
