// Stencil device code shared by the single-grid Yee kernels (dev_yee.hip) and the batched
// multi-geometry kernels (dev_batch.hip): argument blocks, sources, PML terms, monitors, and the
// E / H block bodies.  Reference semantics: timefield3d.c:1679-1837 (semi-implicit E/H updates),
// 3176-3239 (overlap); the PML is the extension whose formulas oracle/fdtd_oracle.c states.
//
// Layout: six FP32 component arrays, cell c = i + I*(j + J*k).  A block body handles one logical
// block: kBlock (i, j) columns of one z chunk; each thread marches its column upward, keeping the k-1
// (E update) or k+1 (H update) neighbours it needs in registers; the i-1/j-1 (i+1/j+1) neighbours
// come through L1/L2.
#pragma once

#include <algorithm>
#include <type_traits>

#include "dev_internal.h"

namespace fdtd {

// Arguments of one split E / H launch over one grid (also one entry of a batch table).
struct YeeArgs {
    float *Ex, *Ey, *Ez, *Hx, *Hy, *Hz;
    const void *mat;
    const float *coef;  // [nmat][kCoefStride] for the chosen scheme
    const uint16_t *ext;  // extruded ids [K][I] of rows [ext_j0, ext_j1), ext_bg elsewhere, or null
    int ext_j0, ext_j1, ext_bg;
    int I, J, K;
    long long P;
    int k0, k1;          // z range updated, already clipped to [1, K-1)
    int kchunk;          // z planes per block
    int plane_blocks;    // blocks per z plane
    int nsrc;
    SrcPlaneArg src[kMaxSrcPlanes];
    const float *dense;  // 6 x N source volume when sources span more than kMaxSrcPlanes planes
    long long N;
    int nrs;
    RunSrcArg rs[kMaxRunSrc];
    int nmon;
    RunMonArg mon[kMaxRunMon];
    // device step counter of graph-captured batch steps (null: per-step values are in the args)
    const int *step;
    int step0;
    // z planes [src_k0, src_k1) may carry a source (dense volume, stored planes or run sources);
    // set by set_source_range once the source fields are final
    int src_k0, src_k1;
};

struct PmlArgs {
    const float *tab[3];  // per axis: bE, aE, kE, bH, aH, kH (N_axis floats each)
    int nlo[3], nhi[3], W[3];
    float *psiE[3][2];
    float *psiH[3][2];
};

// Host side: the z range of the planes that may carry a source, from a's source fields.
inline void set_source_range(YeeArgs &a) {
    int k0 = a.K, k1 = 0;
    if (a.dense) k0 = 0, k1 = a.K;
    for (int s = 0; s < a.nsrc; s++) k0 = std::min(k0, a.src[s].k), k1 = std::max(k1, a.src[s].k + 1);
    for (int q = 0; q < a.nrs; q++) k0 = std::min(k0, a.rs[q].k), k1 = std::max(k1, a.rs[q].k + 1);
    a.src_k0 = k0 < k1 ? k0 : 0;
    a.src_k1 = k0 < k1 ? k1 : 0;
}

// Fill the launch arguments of the full-range split step of grid d (PML per d's boundary,
// semi-implicit scheme, stored source planes, no run sources or monitors); builds the
// coefficient and PML tables for dt first.
int fill_step_args(fdtd_dev *d, double dt, YeeArgs *a, PmlArgs *pm);

// Batched runs (dev_run.hip).  run_batch_prepare readies steps n0..n0+n-1 of grid d's run for
// launches that read the step index from the device counter `step` (device source-coefficient
// tables, monitor descriptors and partial buffers) and fills d's launch arguments with them.
// run_batch_finish, called once the steps of all ng grids are queued on st, forms every
// monitor's per-step overlaps and running DFTs of those steps on st (the twiddles are computed on
// the host meanwhile and go up with one copy from the stage's pinned buffer, which must not be
// reused before st has finished) and leaves each source plane as the last step wrote it.
struct BatchStage;
BatchStage *batch_stage_new();
void batch_stage_free(BatchStage *s);
int run_batch_prepare(fdtd_dev *d, double dt, int n0, int n, const int *step, YeeArgs *a, PmlArgs *pm);
int run_batch_finish(fdtd_dev *const *grids, int ng, double dt, int n0, int n, hipStream_t st, BatchStage *s);

namespace {

// Global-memory view of a device pointer.  A pointer that a kernel loads from memory rather than
// receives as an argument (the batched kernels' per-grid descriptor table) is generic to the
// compiler, which then emits flat memory instructions for it: they need a 64-bit vector address
// per access (no scalar base), and they count against the scalar-load counter too, so each wait
// for a descriptor field also drains the field loads in flight.  Every pointer the helpers below
// take points into device memory, so they access it through this view.
#define FDTD_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ FDTD_GLOBAL T *gview(T *p) {
    return (FDTD_GLOBAL T *)p;
}

// Cell offsets inside one grid's arrays.  Off = uint32_t when every array of the grid is below
// 4 GiB (fewer than 2^30 cells): a load is then a uniform 64-bit base in scalar registers plus a
// 32-bit byte offset in one vector register, and fixed neighbour deltas (c - 1, c - I) become
// the instruction's immediate offset, which removes the per-access 64-bit address arithmetic.
// Off = long long covers larger grids.
template <typename Off, typename T>
__device__ __forceinline__ T ld(const T *base, Off i) {
    if constexpr (sizeof(Off) == 4)
        return *reinterpret_cast<const FDTD_GLOBAL T *>(reinterpret_cast<const FDTD_GLOBAL char *>(gview(base)) +
                                                         (uint32_t)i * (uint32_t)sizeof(T));
    else
        return gview(base)[i];
}

template <typename Off, typename T>
__device__ __forceinline__ void st(T *base, Off i, T v) {
    // streaming stores: the updated fields are read again only by the next launch
    if constexpr (sizeof(Off) == 4)
        __builtin_nontemporal_store(v, reinterpret_cast<FDTD_GLOBAL T *>(reinterpret_cast<FDTD_GLOBAL char *>(gview(base)) +
                                                                          (uint32_t)i * (uint32_t)sizeof(T)));
    else
        __builtin_nontemporal_store(v, gview(base) + i);
}

// Load of a value no other thread of the launch reads (a cell's own old E in the E update, own
// old H in the H update).  NT: with the streaming (non-temporal) hint, for grids whose fields do
// not fit the last-level cache, where keeping these lines only evicts the neighbour rows other
// threads still need (12.4 M-cell taper: E update 105 -> 88 us); in-cache grids read them
// plainly (the hint costs the strip waveguide 20 %).
template <bool NT, typename Off, typename T>
__device__ __forceinline__ T ld_own(const T *base, Off i) {
    if constexpr (!NT) {
        return ld(base, i);
    } else if constexpr (sizeof(Off) == 4) {
        return __builtin_nontemporal_load(reinterpret_cast<const FDTD_GLOBAL T *>(
            reinterpret_cast<const FDTD_GLOBAL char *>(gview(base)) + (uint32_t)i * (uint32_t)sizeof(T)));
    } else {
        return __builtin_nontemporal_load(gview(base) + i);
    }
}

template <typename MatT, typename Off>
__device__ __forceinline__ int mat_at(const void *mat, Off c) {
    return (int)ld(static_cast<const MatT *>(mat), c);
}

// Material ids taken from an extruded geometry's K x I table (YeeArgs::ext) instead of the
// per-cell volume: the batched kernels' MatT when every grid of the batch is extruded.  The table
// of one z plane is a few hundred bytes that neighbouring blocks share through the cache, so a
// cell-update no longer reads its 1- or 2-byte id from HBM.
struct ExtIds {};

// Material id of cell c = (i, j, k).
template <typename MatT, typename Off>
__device__ __forceinline__ int material(const YeeArgs &a, Off c, int i, int j, int k) {
    if constexpr (std::is_same<MatT, ExtIds>::value)
        return j >= a.ext_j0 && j < a.ext_j1 ? (int)gview(a.ext)[k * a.I + i] : a.ext_bg;
    else
        return mat_at<MatT>(a.mat, c);
}

// Source components m0..m0+2 of plane cell p at plane k: tabulated run source inside its mode
// window, else the stored source plane.  Returns false when plane k carries no source.  The
// per-component planes are addressed from uniform bases (the dense 6 x N source volume in 64-bit
// arithmetic), so only the in-plane offset p is per lane.
template <typename Off>
__device__ __forceinline__ bool source3(const YeeArgs &a, int k, Off p, int m0, float v[3]) {
    if (k < a.src_k0 || k >= a.src_k1) return false;  // most planes: two scalar compares
    for (int q = 0; q < a.nrs; q++) {
        const RunSrcArg &r = a.rs[q];
        if (r.k == k && ld(r.win, p)) {
            float c1 = r.c1, c2 = r.c2;
            if (r.tab) {
                const int s = *gview(a.step);
                const FDTD_GLOBAL float *t = gview(reinterpret_cast<const float *>(r.tab));  // (c1, c2) pairs
                const bool in = s >= 0 && s < r.ntab;
                c1 = in ? t[2 * s] : 0.0f;
                c2 = in ? t[2 * s + 1] : 0.0f;
            }
            for (int m = 0; m < 3; m++)
                v[m] = c1 * ld(r.PR + (long long)(m0 + m) * a.P, p) + c2 * ld(r.PI + (long long)(m0 + m) * a.P, p);
            return true;
        }
    }
    const float *J = nullptr;
    long long stride = a.P;
    if (a.dense) {
        J = a.dense + (long long)k * a.P;
        stride = a.N;
    } else {
        for (int s = 0; s < a.nsrc; s++)
            if (a.src[s].k == k) J = a.src[s].J;
    }
    if (!J) return false;
    for (int m = 0; m < 3; m++) v[m] = ld(J + (long long)(m0 + m) * stride, p);
    return true;
}

__device__ __forceinline__ int slab_index(int n, int N, int nlo, int nhi) {
    if (n < nlo) return n;
    if (nhi > 0 && n >= N - 1 - nhi) return nlo + n - (N - 1 - nhi);
    return -1;
}

// psi <- b psi + a d from the psi value `old` loaded earlier (stored only when `live`); returns the
// PML-corrected difference d/kappa + psi
template <typename Off>
__device__ __forceinline__ float pml_term(float *psi, Off q, float old, float b, float a, float kinv, float d,
                                          bool live) {
    float v = b * old + a * d;
    if (live) st(psi, q, v);
    return d * kinv + v;
}

// psi offsets of cell (i, j, k) in the x-, y- and z-slab arrays (meaningful only inside that slab)
template <typename Off>
__device__ __forceinline__ Off psi_x(const PmlArgs &pm, const YeeArgs &a, int j, int k, int sx) {
    return (Off)sx + (Off)pm.W[0] * ((Off)j + (Off)a.J * (Off)k);
}
template <typename Off>
__device__ __forceinline__ Off psi_y(const PmlArgs &pm, const YeeArgs &a, int i, int k, int sy) {
    return (Off)i + (Off)a.I * ((Off)sy + (Off)pm.W[1] * (Off)k);
}
template <typename Off>
__device__ __forceinline__ Off psi_z(const YeeArgs &a, Off p, int sz) {
    return p + (Off)a.P * (Off)sz;
}

// The six psi values of cell (i, j, k) (two per PML axis the cell lies in; 0 for the others),
// loaded together with the cell's field inputs so that their latency overlaps the fields' instead
// of following it.  psi = psiE or psiH of the PML arguments.
template <typename Off>
__device__ __forceinline__ void psi_load(float *const psi[3][2], const PmlArgs &pm, const YeeArgs &a, int i, int j,
                                         Off p, int k, int sx, int sy, float v[6]) {
    v[0] = v[1] = v[2] = v[3] = v[4] = v[5] = 0.0f;
    const int sz = slab_index(k, a.K, pm.nlo[2], pm.nhi[2]);
    if (sx >= 0) {
        const Off q = psi_x<Off>(pm, a, j, k, sx);
        v[0] = ld(psi[0][0], q);
        v[1] = ld(psi[0][1], q);
    }
    if (sy >= 0) {
        const Off q = psi_y<Off>(pm, a, i, k, sy);
        v[2] = ld(psi[1][0], q);
        v[3] = ld(psi[1][1], q);
    }
    if (sz >= 0) {
        const Off q = psi_z<Off>(a, p, sz);
        v[4] = ld(psi[2][0], q);
        v[5] = ld(psi[2][1], q);
    }
}

// E <- caE*E + cbE*(backward-difference curl of H) - cjE*J   (timefield3d.c:1759-1837)
// KC > 0: the z chunk is KC planes, all loaded before any is computed (more bytes in flight per
// wavefront); iterations past the end of a short last chunk reload its last plane and store
// nothing.  KC = 0: a plain loop over a runtime chunk.  Inside the PML the differences take the
// CPML correction d/kappa + psi; a cell outside every slab has kappa = 1 and no psi on all axes
// (the tables hold 1, 0, 0 there), so it skips the correction altogether.
template <typename MatT, bool PML, int KC, typename Off, bool NT = false>
__device__ __forceinline__ void update_E_block(const YeeArgs &a, const PmlArgs &pm, int lb) {
    const int pb = lb % a.plane_blocks, kc = lb / a.plane_blocks;
    const Off p = (Off)pb * (Off)kBlock + (Off)threadIdx.x;
    if (p >= (Off)a.P) return;
    const int j = (int)(p / (Off)a.I), i = (int)(p - (Off)j * (Off)a.I);
    if (i < 1 || i > a.I - 2 || j < 1 || j > a.J - 2) return;
    const int kb = a.k0 + kc * a.kchunk;
    const int ke = min(a.k1, kb + a.kchunk);
    if (kb >= ke) return;
    float *__restrict__ Ex = a.Ex;
    float *__restrict__ Ey = a.Ey;
    float *__restrict__ Ez = a.Ez;
    const float *__restrict__ Hx = a.Hx;
    const float *__restrict__ Hy = a.Hy;
    const float *__restrict__ Hz = a.Hz;
    const Off P = (Off)a.P, I = (Off)a.I;
    int sx = -1, sy = -1;
    float bx = 0, ax = 0, kx = 1, by = 0, ay = 0, ky = 1;
    if (PML) {
        sx = slab_index(i, a.I, pm.nlo[0], pm.nhi[0]);
        sy = slab_index(j, a.J, pm.nlo[1], pm.nhi[1]);
        if (sx >= 0) bx = gview(pm.tab[0])[i], ax = gview(pm.tab[0])[a.I + i], kx = gview(pm.tab[0])[2 * a.I + i];
        if (sy >= 0) by = gview(pm.tab[1])[j], ay = gview(pm.tab[1])[a.J + j], ky = gview(pm.tab[1])[2 * a.J + j];
    }
    // one cell-update of plane k from its loaded inputs; stores only when `live`
    auto cell = [&](int k, bool live, Off c, float hx, float hy, float hz, float hy_im, float hz_im, float hx_jm,
                    float hz_jm, float hx_km, float hy_km, float eox, float eoy, float eoz, int m, const float *pv) {
        const FDTD_GLOBAL float *cf = gview(a.coef) + m * kCoefStride;
        // curl_x = dHz/dy - dHy/dz, curl_y = dHx/dz - dHz/dx, curl_z = dHy/dx - dHx/dy
        float x1 = hz - hz_jm, x2 = hy - hy_km;
        float y1 = hx - hx_km, y2 = hz - hz_im;
        float z1 = hy - hy_im, z2 = hx - hx_jm;
        const int sz = PML ? slab_index(k, a.K, pm.nlo[2], pm.nhi[2]) : -1;
        if (PML && (sx >= 0 || sy >= 0 || sz >= 0)) {
            const float kz = sz >= 0 ? gview(pm.tab[2])[2 * a.K + k] : 1.0f;
            float tx1 = x1 * ky, tx2 = x2 * kz, ty1 = y1 * kz, ty2 = y2 * kx, tz1 = z1 * kx, tz2 = z2 * ky;
            if (sx >= 0) {
                const Off q = psi_x<Off>(pm, a, j, k, sx);
                ty2 = pml_term(pm.psiE[0][0], q, pv[0], bx, ax, kx, y2, live);
                tz1 = pml_term(pm.psiE[0][1], q, pv[1], bx, ax, kx, z1, live);
            }
            if (sy >= 0) {
                const Off q = psi_y<Off>(pm, a, i, k, sy);
                tz2 = pml_term(pm.psiE[1][0], q, pv[2], by, ay, ky, z2, live);
                tx1 = pml_term(pm.psiE[1][1], q, pv[3], by, ay, ky, x1, live);
            }
            if (sz >= 0) {
                const float bz = gview(pm.tab[2])[k], az = gview(pm.tab[2])[a.K + k];
                const Off q = psi_z<Off>(a, p, sz);
                tx2 = pml_term(pm.psiE[2][0], q, pv[4], bz, az, kz, x2, live);
                ty1 = pml_term(pm.psiE[2][1], q, pv[5], bz, az, kz, y1, live);
            }
            x1 = tx1, x2 = tx2, y1 = ty1, y2 = ty2, z1 = tz1, z2 = tz2;
        }
        float ex = cf[0] * eox + cf[3] * (x1 - x2);
        float ey = cf[1] * eoy + cf[4] * (y1 - y2);
        float ez = cf[2] * eoz + cf[5] * (z1 - z2);
        float sv[3];
        if (source3(a, k, p, 0, sv)) {
            ex -= cf[6] * sv[0];
            ey -= cf[7] * sv[1];
            ez -= cf[8] * sv[2];
        }
        if (live) {
            st(Ex, c, ex);
            st(Ey, c, ey);
            st(Ez, c, ez);
        }
    };
    const Off c0 = p + (Off)kb * P;
    if constexpr (KC > 0) {
        float hx[KC], hy[KC], hz[KC], hy_im[KC], hz_im[KC], hx_jm[KC], hz_jm[KC], eox[KC], eoy[KC], eoz[KC];
        float pv[KC][6];
        int m[KC];
        const float hx_k0 = ld(Hx, c0 - P), hy_k0 = ld(Hy, c0 - P);
#pragma unroll
        for (int t = 0; t < KC; t++) {
            const int kk = kb + t < ke ? kb + t : ke - 1;
            const Off c = p + (Off)kk * P;
            hx[t] = ld(Hx, c), hy[t] = ld(Hy, c), hz[t] = ld(Hz, c);
            hy_im[t] = ld(Hy, c - 1), hz_im[t] = ld(Hz, c - 1);
            hx_jm[t] = ld(Hx, c - I), hz_jm[t] = ld(Hz, c - I);
            eox[t] = ld_own<NT>(Ex, c), eoy[t] = ld_own<NT>(Ey, c), eoz[t] = ld_own<NT>(Ez, c);
            m[t] = material<MatT>(a, c, i, j, kk);
            if (PML) psi_load(pm.psiE, pm, a, i, j, p, kk, sx, sy, pv[t]);
        }
#pragma unroll
        for (int t = 0; t < KC; t++) {
            const bool live = kb + t < ke;
            const int kk = live ? kb + t : ke - 1;
            cell(kk, live, p + (Off)kk * P, hx[t], hy[t], hz[t], hy_im[t], hz_im[t], hx_jm[t], hz_jm[t],
                 t == 0 ? hx_k0 : hx[t - 1], t == 0 ? hy_k0 : hy[t - 1], eox[t], eoy[t], eoz[t], m[t], pv[t]);
        }
    } else {
        float hx_km = ld(Hx, c0 - P), hy_km = ld(Hy, c0 - P);
        Off c = c0;
        for (int k = kb; k < ke; k++, c += P) {
            const float hx = ld(Hx, c), hy = ld(Hy, c), hz = ld(Hz, c);
            float pv[6];
            if (PML) psi_load(pm.psiE, pm, a, i, j, p, k, sx, sy, pv);
            cell(k, true, c, hx, hy, hz, ld(Hy, c - 1), ld(Hz, c - 1), ld(Hx, c - I), ld(Hz, c - I), hx_km, hy_km,
                 ld_own<NT>(Ex, c), ld_own<NT>(Ey, c), ld_own<NT>(Ez, c), material<MatT>(a, c, i, j, k), pv);
            hx_km = hx;
            hy_km = hy;
        }
    }
}

// Block-wide FP64 sum of (re, im); every thread of the block must call it.  Returns the sum in
// all threads.
__device__ __forceinline__ double2 block_sum2(double re, double im) {
    __shared__ double sre[kBlock / 64], sim[kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) {
        re += __shfl_down(re, off, 64);
        im += __shfl_down(im, off, 64);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sre[w] = re;
        sim[w] = im;
    }
    __syncthreads();
    double r = 0, s = 0;
    for (int q = 0; q < kBlock / 64; q++) {
        r += sre[q];
        s += sim[q];
    }
    __syncthreads();  // the slots are reused by the next call
    return make_double2(r, s);
}

// Overlap monitors of a MON launch, after the block's z march: for each monitor plane inside the
// chunk, the block's FP64 partial of Hx mEy - Hy mEx + Ey mHx - Ex mHy (timefield3d.c:3195-3224)
// over its cells, from E (final since the E update) and H (just stored by these same threads).
// Ex, Ey, Hx, Hy: the arrays holding the step's final fields (the launch's own, or the second
// buffer of a ping-pong step).
template <typename Off>
__device__ __forceinline__ void monitor_epilogue_on(const YeeArgs &a, const float *Ex, const float *Ey, const float *Hx,
                                                    const float *Hy, int kb, int ke, int pb, Off p, bool active) {
    for (int q = 0; q < a.nmon; q++) {
        const RunMonArg &M = a.mon[q];
        if (M.k < kb || M.k >= ke) continue;  // uniform over the block
        double re = 0.0, im = 0.0;
        if (active) {
            const Off c = p + (Off)M.k * (Off)a.P;
            const double ex = ld(Ex, c), ey = ld(Ey, c), hx = ld(Hx, c), hy = ld(Hy, c);
            double w[8];
            for (int r = 0; r < 8; r++) w[r] = ld(M.W + (long long)r * a.P, p);
            re = (hx * w[2] - hy * w[0]) + (ey * w[4] - ex * w[6]);
            im = (hx * w[3] - hy * w[1]) + (ey * w[5] - ex * w[7]);
        }
        const double2 v = block_sum2(re, im);
        double2 *part = a.step ? M.part + (long long)(*gview(a.step) - a.step0) * M.stride : M.part;
        if (threadIdx.x == 0) {
            FDTD_GLOBAL double *q = gview(reinterpret_cast<double *>(part + pb));
            q[0] = v.x;
            q[1] = v.y;
        }
    }
}

template <typename Off>
__device__ __forceinline__ void monitor_epilogue(const YeeArgs &a, int kb, int ke, int pb, Off p, bool active) {
    monitor_epilogue_on(a, a.Ex, a.Ey, a.Hx, a.Hy, kb, ke, pb, p, active);
}

// H <- caH*H - cbH*(forward-difference curl of E) - cjH*M   (timefield3d.c:1679-1757)
// KC and the PML treatment as for the E update.  MON: the launch also evaluates the run's overlap
// monitors (every thread of a block reaches the block reductions, so out-of-range threads skip
// the march instead of returning).
template <typename MatT, bool PML, bool MON, int KC, typename Off, bool NT = false>
__device__ __forceinline__ void update_H_block(const YeeArgs &a, const PmlArgs &pm, int lb) {
    const int pb = lb % a.plane_blocks, kc = lb / a.plane_blocks;
    const Off p = (Off)pb * (Off)kBlock + (Off)threadIdx.x;
    const int kb = a.k0 + kc * a.kchunk;
    const int ke = min(a.k1, kb + a.kchunk);
    if (kb >= ke) return;  // uniform over the block
    const int j = (int)(p / (Off)a.I), i = (int)(p - (Off)j * (Off)a.I);
    const bool active = p < (Off)a.P && i >= 1 && i <= a.I - 2 && j >= 1 && j <= a.J - 2;
    if (!MON && !active) return;
    if (active) {
        const float *__restrict__ Ex = a.Ex;
        const float *__restrict__ Ey = a.Ey;
        const float *__restrict__ Ez = a.Ez;
        float *__restrict__ Hx = a.Hx;
        float *__restrict__ Hy = a.Hy;
        float *__restrict__ Hz = a.Hz;
        const Off P = (Off)a.P, I = (Off)a.I;
        int sx = -1, sy = -1;
        float bx = 0, ax = 0, kx = 1, by = 0, ay = 0, ky = 1;
        if (PML) {
            sx = slab_index(i, a.I, pm.nlo[0], pm.nhi[0]);
            sy = slab_index(j, a.J, pm.nlo[1], pm.nhi[1]);
            if (sx >= 0) bx = gview(pm.tab[0])[3 * a.I + i], ax = gview(pm.tab[0])[4 * a.I + i], kx = gview(pm.tab[0])[5 * a.I + i];
            if (sy >= 0) by = gview(pm.tab[1])[3 * a.J + j], ay = gview(pm.tab[1])[4 * a.J + j], ky = gview(pm.tab[1])[5 * a.J + j];
        }
        auto cell = [&](int k, bool live, Off c, float ex, float ey, float ez, float ex_kp, float ey_kp, float ey_ip,
                        float ez_ip, float ex_jp, float ez_jp, float hox, float hoy, float hoz, int m, const float *pv) {
            const FDTD_GLOBAL float *cf = gview(a.coef) + m * kCoefStride;
            // curl_x = dEz/dy - dEy/dz, curl_y = dEx/dz - dEz/dx, curl_z = dEy/dx - dEx/dy
            float x1 = ez_jp - ez, x2 = ey_kp - ey;
            float y1 = ex_kp - ex, y2 = ez_ip - ez;
            float z1 = ey_ip - ey, z2 = ex_jp - ex;
            const int sz = PML ? slab_index(k, a.K, pm.nlo[2], pm.nhi[2]) : -1;
            if (PML && (sx >= 0 || sy >= 0 || sz >= 0)) {
                const float kz = sz >= 0 ? gview(pm.tab[2])[5 * a.K + k] : 1.0f;
                float tx1 = x1 * ky, tx2 = x2 * kz, ty1 = y1 * kz, ty2 = y2 * kx, tz1 = z1 * kx, tz2 = z2 * ky;
                if (sx >= 0) {
                    const Off q = psi_x<Off>(pm, a, j, k, sx);
                    ty2 = pml_term(pm.psiH[0][0], q, pv[0], bx, ax, kx, y2, live);
                    tz1 = pml_term(pm.psiH[0][1], q, pv[1], bx, ax, kx, z1, live);
                }
                if (sy >= 0) {
                    const Off q = psi_y<Off>(pm, a, i, k, sy);
                    tz2 = pml_term(pm.psiH[1][0], q, pv[2], by, ay, ky, z2, live);
                    tx1 = pml_term(pm.psiH[1][1], q, pv[3], by, ay, ky, x1, live);
                }
                if (sz >= 0) {
                    const float bz = gview(pm.tab[2])[3 * a.K + k], az = gview(pm.tab[2])[4 * a.K + k];
                    const Off q = psi_z<Off>(a, p, sz);
                    tx2 = pml_term(pm.psiH[2][0], q, pv[4], bz, az, kz, x2, live);
                    ty1 = pml_term(pm.psiH[2][1], q, pv[5], bz, az, kz, y1, live);
                }
                x1 = tx1, x2 = tx2, y1 = ty1, y2 = ty2, z1 = tz1, z2 = tz2;
            }
            float hx = cf[9] * hox - cf[12] * (x1 - x2);
            float hy = cf[10] * hoy - cf[13] * (y1 - y2);
            float hz = cf[11] * hoz - cf[14] * (z1 - z2);
            float sv[3];
            if (source3(a, k, p, 3, sv)) {
                hx -= cf[15] * sv[0];
                hy -= cf[16] * sv[1];
                hz -= cf[17] * sv[2];
            }
            if (live) {
                st(Hx, c, hx);
                st(Hy, c, hy);
                st(Hz, c, hz);
            }
        };
        const Off c0 = p + (Off)kb * P;
        if constexpr (KC > 0) {
            float ez[KC], ex_kp[KC], ey_kp[KC], ey_ip[KC], ez_ip[KC], ex_jp[KC], ez_jp[KC], hox[KC], hoy[KC], hoz[KC];
            float pv[KC][6];
            int m[KC];
            const float ex_k0 = ld(Ex, c0), ey_k0 = ld(Ey, c0);
#pragma unroll
            for (int t = 0; t < KC; t++) {
                const int kk = kb + t < ke ? kb + t : ke - 1;
                const Off c = p + (Off)kk * P;
                ez[t] = ld(Ez, c);
                ex_kp[t] = ld(Ex, c + P), ey_kp[t] = ld(Ey, c + P);
                ey_ip[t] = ld(Ey, c + 1), ez_ip[t] = ld(Ez, c + 1);
                ex_jp[t] = ld(Ex, c + I), ez_jp[t] = ld(Ez, c + I);
                hox[t] = ld_own<NT>(Hx, c), hoy[t] = ld_own<NT>(Hy, c), hoz[t] = ld_own<NT>(Hz, c);
                m[t] = material<MatT>(a, c, i, j, kk);
                if (PML) psi_load(pm.psiH, pm, a, i, j, p, kk, sx, sy, pv[t]);
            }
#pragma unroll
            for (int t = 0; t < KC; t++) {
                const bool live = kb + t < ke;
                const int kk = live ? kb + t : ke - 1;
                cell(kk, live, p + (Off)kk * P, t == 0 ? ex_k0 : ex_kp[t - 1], t == 0 ? ey_k0 : ey_kp[t - 1], ez[t],
                     ex_kp[t], ey_kp[t], ey_ip[t], ez_ip[t], ex_jp[t], ez_jp[t], hox[t], hoy[t], hoz[t], m[t], pv[t]);
            }
        } else {
            float ex = ld(Ex, c0), ey = ld(Ey, c0);
            Off c = c0;
            for (int k = kb; k < ke; k++, c += P) {
                const float ex_kp = ld(Ex, c + P), ey_kp = ld(Ey, c + P);
                float pv[6];
                if (PML) psi_load(pm.psiH, pm, a, i, j, p, k, sx, sy, pv);
                cell(k, true, c, ex, ey, ld(Ez, c), ex_kp, ey_kp, ld(Ey, c + 1), ld(Ez, c + 1), ld(Ex, c + I),
                     ld(Ez, c + I), ld_own<NT>(Hx, c), ld_own<NT>(Hy, c), ld_own<NT>(Hz, c), material<MatT>(a, c, i, j, k),
                     pv);
                ex = ex_kp;
                ey = ey_kp;
            }
        }
    }
    if (MON) monitor_epilogue(a, kb, ke, pb, p, active);
}

}  // namespace
}  // namespace fdtd
