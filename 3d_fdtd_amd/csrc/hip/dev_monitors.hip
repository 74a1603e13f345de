// Sources, monitors and adjoint histories on gfx950.
//   mode sources           timefield3d.c:2977-3068
//   single-frequency DFT   timefield3d.c:2851-2975   (FP64 accumulators)
//   mode overlap           timefield3d.c:3176-3239   (FP64 two-stage deterministic reduction)
//   taper edge store/grad  adjoint3d.c:259-373
//   3-D block adjoint      adjoint3d.c:48-105
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>

#include "dev_internal.h"

namespace fdtd {

// Scratch of the allocation-free overlap (fdtd_dev_mode_overlap_noalloc): kNoAllocPartials block
// partials, then the last-block counter in the first word of the final slot.  Each GPU holds its
// own copy until the process ends, and every call on that GPU shares it.  Host code never passes
// this name to a kernel: it stands for the host-side shadow, not the device address, which
// noalloc_scratch() looks up once per device.
constexpr int kNoAllocPartials = 8192;  // planes of up to 8192 * 128 cells
static __device__ __align__(256) double2 scratch[kNoAllocPartials + 1];

namespace {

std::map<const void *, TaperHist> g_tapers;
std::map<const void *, BlockHist> g_blocks;

template <typename MatT>
__device__ __forceinline__ int mat_id(const void *mat, long long c) {
    return (int)static_cast<const MatT *>(mat)[c];
}

__device__ __forceinline__ int mode_index_1d(int n, int center, double ratio, int mn) {
    double dist = (n - center) * ratio;
    return (int)dist + mn / 2;
}

// Forward soft source: J = A ep Re(mode e^{-i phase}), M = -(2r-1) A mu Re(...), stored as float;
// adjoint CW source: J = (Re Hy, -Re Hx, 0), M = (-(2r-1) Re Ey, (2r-1) Re Ex, 0).
// Cells outside the mode window keep their previous value.
template <typename MatT>
__global__ void k_mode_source(float *plane, long long cs, const void *mat, const float *props, int I, int J,
                              long long P, int cz,
                              int cx, int cy, const double *mode, int mi, int mj, double ratio, double cp, double sp,
                              double amp, int msign, int adjoint) {
#pragma clang fp contract(off)
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int j = (int)(p / I), i = (int)(p - (long long)j * I);
    const int mx = mode_index_1d(i, cx, ratio, mi), my = mode_index_1d(j, cy, ratio, mj);
    if (mx < 0 || mx >= mi || my < 0 || my >= mj) return;
    const double *md = mode + 12 * (mx + (long long)mi * my);
    if (!adjoint) {
        const float *pr = props + (long long)mat_id<MatT>(mat, p + (long long)cz * P) * kPropStride;
        for (int m = 0; m < 6; m++) {
            double proj = md[2 * m] * cp + md[2 * m + 1] * sp;
            double v;
            if (m < 3) v = amp * (double)pr[m] * proj;
            else v = (double)(-msign) * amp * (double)pr[3 + m % 3] * proj;
            plane[m * cs + p] = (float)v;
        }
    } else {
        double rEx = md[0] * cp + md[1] * sp;
        double rEy = md[2] * cp + md[3] * sp;
        double rHx = md[6] * cp + md[7] * sp;
        double rHy = md[8] * cp + md[9] * sp;
        plane[0 * cs + p] = (float)rHy;
        plane[1 * cs + p] = (float)(-rHx);
        plane[2 * cs + p] = 0.0f;
        plane[3 * cs + p] = (float)((double)(-1 * msign) * rEy);
        plane[4 * cs + p] = (float)((double)msign * rEx);
        plane[5 * cs + p] = 0.0f;
    }
}

__global__ void k_plane_masked_copy(float *plane, long long cs, const float *src, const unsigned char *mask,
                                    long long P) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    if (mask && !mask[p]) return;
    for (int m = 0; m < 6; m++) plane[m * cs + p] = src[m * P + p];
}

// Overlap stage 1 (one block): the block's FP64 partial sum of Hx mEy - Hy mEx + Ey mHx - Ex mHy,
// stored in partials[blockIdx.x] by thread 0.
__device__ __forceinline__ void overlap_block_partial(const float *Ex, const float *Ey, const float *Hx,
                                                      const float *Hy, int I, int J, long long P, int cz, int cx,
                                                      int cy, const double *mode, int mi, int mj, double ratio,
                                                      int reverse, double2 *partials) {
    __shared__ double sre[kBlock / 64], sim[kBlock / 64];
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double re = 0.0, im = 0.0;
    if (p < P) {
        const int j = (int)(p / I), i = (int)(p - (long long)j * I);
        const int mx = mode_index_1d(i, cx, ratio, mi), my = mode_index_1d(j, cy, ratio, mj);
        if (mx >= 0 && mx < mi && my >= 0 && my < mj) {
            const double *md = mode + 12 * (mx + (long long)mi * my);
            double mEx_r = md[0], mEx_i = md[1], mEy_r = md[2], mEy_i = md[3];
            double mHx_r = md[6], mHx_i = md[7], mHy_r = md[8], mHy_i = md[9];
            if (reverse == 1) {  // mH -> -conj(mH)
                mHx_r = -mHx_r;
                mHy_r = -mHy_r;
            }
            const long long c = p + (long long)cz * P;
            const double ex = Ex[c], ey = Ey[c], hx = Hx[c], hy = Hy[c];
            re = (hx * mEy_r - hy * mEx_r) + (ey * mHx_r - ex * mHy_r);
            im = (hx * mEy_i - hy * mEx_i) + (ey * mHx_i - ex * mHy_i);
        }
    }
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
    if (threadIdx.x == 0) {
        double r = 0, s = 0;
        for (int q = 0; q < kBlock / 64; q++) {
            r += sre[q];
            s += sim[q];
        }
        partials[blockIdx.x] = make_double2(r, s);
    }
}

// Overlap stage 2 (one block): sums the n partials in a fixed order and writes scale * sum.
__device__ __forceinline__ void overlap_block_final(const double2 *partials, int n, double scale, double2 *out) {
    __shared__ double sre[kBlock], sim[kBlock];
    double r = 0, s = 0;
    for (int q = threadIdx.x; q < n; q += kBlock) {
        r += partials[q].x;
        s += partials[q].y;
    }
    sre[threadIdx.x] = r;
    sim[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sre[threadIdx.x] += sre[threadIdx.x + w];
            sim[threadIdx.x] += sim[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = make_double2(scale * sre[0], scale * sim[0]);
}

// The plane overlap in one launch: every block stores its partial; the last block to finish (a
// counter in device memory, reset by that block for the next launch) sums them in a fixed order,
// so the result does not depend on which block finishes last.  `out` may be host memory.
__global__ __launch_bounds__(kBlock) void k_overlap(const float *Ex, const float *Ey, const float *Hx,
                                                    const float *Hy, int I, int J, long long P, int cz, int cx,
                                                    int cy, const double *mode, int mi, int mj, double ratio,
                                                    int reverse, double2 *partials, unsigned *count, double scale,
                                                    double2 *out) {
    overlap_block_partial(Ex, Ey, Hx, Hy, I, J, P, cz, cx, cy, mode, mi, mj, ratio, reverse, partials);
    __shared__ bool last;
    if (threadIdx.x == 0) {
        __threadfence();  // this block's partial before its count
        last = atomicAdd(count, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();  // every other block's partial after their counts
    overlap_block_final(partials, (int)gridDim.x, scale, out);
    if (threadIdx.x == 0) {
        *count = 0u;
        __threadfence_system();
    }
}

__global__ void k_clear_counter(unsigned *count) {
    if (threadIdx.x == 0) *count = 0u;
}

// Device address of this GPU's scratch array, looked up (hipGetSymbolAddress) on the first call
// on each device, with that device current; the first call also zeroes the counter with a kernel
// on `st` (every overlap's last block leaves it at zero again).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double2 *g_scratch_ptr[kMaxDevices] = {};

int noalloc_scratch(hipStream_t st, double2 **out) {
    int dev = 0;
    FDTD_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("overlap (no-alloc): device index beyond the scratch table");
        return -2;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) return fail("hipGetSymbolAddress(scratch)", e != hipSuccess ? e : hipErrorUnknown);
        auto *p = static_cast<double2 *>(ptr);
        hipLaunchKernelGGL(k_clear_counter, dim3(1), dim3(64), 0, st, reinterpret_cast<unsigned *>(p + kNoAllocPartials));
        FDTD_TRY(hipGetLastError());
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// F += f * twiddle over a z range, optionally one y row (timefield3d.c:2851-2975).
__global__ void k_dft(double2 *F, const float *base, long long Nal, long long N, int I, int J, long long P, int k0,
                      int k1, int yslice, double twr, double twi) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long c;
    if (yslice >= 0) {
        long long n = (long long)I * (k1 - k0);
        if (t >= n) return;
        int k = k0 + (int)(t / I), i = (int)(t % I);
        c = i + (long long)I * (yslice + (long long)J * k);
    } else {
        long long n = P * (k1 - k0);
        if (t >= n) return;
        c = t + (long long)k0 * P;
    }
#pragma unroll
    for (int m = 0; m < 6; m++) {
        double v = base[m * Nal + c];
        double2 a = F[m * N + c];
        a.x = a.x + v * twr;
        a.y = a.y + v * twi;
        F[m * N + c] = a;
    }
}

// Gather the six components at the right (side 0) and left (side 1) edge pixels of every
// taper row into hist[t][k][j][side][m] (adjoint3d.c:259-337).
__global__ void k_taper_gather(float *hist, const float *base, long long Nal, const int *edge_i, int I, int J,
                               int start, int L, int bottom, int H, int t) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 2 * H * L) return;
    int side = q & 1, jj = (q >> 1) % H, kk = (q >> 1) / H;
    int i = edge_i[side * L + kk];
    long long c = i + (long long)I * ((bottom + jj) + (long long)J * (start + kk));
    float *o = hist + 6LL * (side + 2LL * (jj + (long long)H * (kk + (long long)L * t)));
#pragma unroll
    for (int m = 0; m < 6; m++) o[m] = base[m * Nal + c];
}

// g[side,j,k] = sum_t sum_{m<3} fwd*adj in FP64, t outer and m inner (adjoint3d.c:339-373).
__global__ void k_taper_grad(const float *fwd, const float *adj, int H, int L, int T, double *g) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 2 * H * L) return;
    const long long stride = 12LL * H * L;
    const float *f = fwd + 6LL * q;
    const float *a = adj + 6LL * q;
    double s = 0;
    for (int t = 0; t < T; t++, f += stride, a += stride) {
        for (int m = 0; m < 3; m++) s += (double)f[m] * (double)a[m];
    }
    g[q] = s;
}

__global__ void k_block_gather(float *fwd, const float *base, long long Nal, int I, int J, int x0, int nx, int y0,
                               int ny, int z0, int nz, int t) {
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long n = (long long)nx * ny * nz;
    if (q >= n) return;
    int xi = (int)(q % nx), yi = (int)((q / nx) % ny), zi = (int)(q / ((long long)nx * ny));
    long long c = (x0 + xi) + (long long)I * ((y0 + yi) + (long long)J * (z0 + zi));
    float *o = fwd + 3 * (q + n * t);
    for (int m = 0; m < 3; m++) o[m] = base[m * Nal + c];
}

__global__ void k_block_accumulate(double *grad, const float *fwd, const float *base, long long Nal, int I, int J,
                                   int x0, int nx, int y0, int ny, int z0, int nz, int t) {
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long n = (long long)nx * ny * nz;
    if (q >= n) return;
    int xi = (int)(q % nx), yi = (int)((q / nx) % ny), zi = (int)(q / ((long long)nx * ny));
    long long c = (x0 + xi) + (long long)I * ((y0 + yi) + (long long)J * (z0 + zi));
    const float *f = fwd + 3 * (q + n * t);
    double s = 0;
    for (int m = 0; m < 3; m++) s += (double)f[m] * (double)base[m * Nal + c];
    grad[q] += s;
}

}  // namespace

SrcPlane *plane_for(fdtd_dev *d, int k, int *err) {
    *err = 0;
    if (d->dense_src) {
        d->dense_view.k = k;
        d->dense_view.J = d->dense_src + (long long)k * d->P;
        d->dense_view.stride = d->N;
        return &d->dense_view;
    }
    for (auto &sp : d->src)
        if (sp.k == k) return &sp;
    if ((int)d->src.size() >= kMaxSrcPlanes) {
        set_error("sources on more than 8 distinct z planes are not supported by the device path");
        *err = -3;
        return nullptr;
    }
    SrcPlane sp;
    sp.k = k;
    sp.stride = d->P;
    hipError_t e = dev_malloc(&sp.J, (size_t)6 * d->P * sizeof(float));
    if (e == hipSuccess) e = hipMemsetAsync(sp.J, 0, (size_t)6 * d->P * sizeof(float), d->stream);
    if (e != hipSuccess) {
        *err = fail("source plane", e);
        return nullptr;
    }
    d->src.push_back(sp);
    return &d->src.back();
}

}  // namespace fdtd

using namespace fdtd;

extern "C" {

int fdtd_dev_mode_source(fdtd_dev *d, int cx, int cy, int cz, const double *mode, int mi, int mj, double mdisc,
                         double cos_phase, double sin_phase, double amplitude, int reverse, int adjoint) {
    if (cz < 0 || cz >= d->K) {
        set_error("source plane outside the grid");
        return -2;
    }
    if (!d->have_mats) {
        set_error("materials were never uploaded");
        return -1;
    }
    int err;
    SrcPlane *sp = plane_for(d, cz, &err);
    if (!sp) return err;
    const double *dm = mode_on_device(d, mode, mi, mj, mdisc, &err);
    if (!dm) return err;
    unsigned nb = (unsigned)((d->P + kBlock - 1) / kBlock);
    int msign = -1 + 2 * reverse;
    double ratio = d->disc / mdisc;
    if (d->mat_bytes == 1)
        hipLaunchKernelGGL(k_mode_source<uint8_t>, dim3(nb), dim3(kBlock), 0, d->stream, sp->J, sp->stride, d->mat,
                           d->d_props, d->I, d->J, d->P, cz, cx, cy, dm, mi, mj, ratio, cos_phase, sin_phase,
                           amplitude, msign, adjoint);
    else if (d->mat_bytes == 2)
        hipLaunchKernelGGL(k_mode_source<uint16_t>, dim3(nb), dim3(kBlock), 0, d->stream, sp->J, sp->stride, d->mat,
                           d->d_props, d->I, d->J, d->P, cz, cx, cy, dm, mi, mj, ratio, cos_phase, sin_phase,
                           amplitude, msign, adjoint);
    else
        hipLaunchKernelGGL(k_mode_source<uint32_t>, dim3(nb), dim3(kBlock), 0, d->stream, sp->J, sp->stride, d->mat,
                           d->d_props, d->I, d->J, d->P, cz, cx, cy, dm, mi, mj, ratio, cos_phase, sin_phase,
                           amplitude, msign, adjoint);
    FDTD_TRY(hipGetLastError());
    return 0;
}

int fdtd_dev_set_source_plane(fdtd_dev *d, int k, const float *plane, const unsigned char *mask) {
    if (k < 0 || k >= d->K) {
        set_error("source plane outside the grid");
        return -2;
    }
    int err;
    SrcPlane *sp = plane_for(d, k, &err);
    if (!sp) return err;
    std::vector<float> soa((size_t)6 * d->P);
    for (long long c = 0; c < d->P; c++)
        for (int m = 0; m < 6; m++) soa[m * d->P + c] = plane[6 * c + m];
    float *tmp = nullptr;
    unsigned char *dmask = nullptr;
    FDTD_TRY(dev_malloc(&tmp, soa.size() * sizeof(float)));
    hipError_t e = hipMemcpy(tmp, soa.data(), soa.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess && mask) {
        e = dev_malloc(&dmask, (size_t)d->P);
        if (e == hipSuccess) e = hipMemcpy(dmask, mask, (size_t)d->P, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) {
        unsigned nb = (unsigned)((d->P + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(k_plane_masked_copy, dim3(nb), dim3(kBlock), 0, d->stream, sp->J, sp->stride, tmp, dmask,
                           d->P);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    (void)dev_free(tmp);
    (void)dev_free(dmask);
    if (e != hipSuccess) return fail("fdtd_dev_set_source_plane", e);
    return 0;
}

int fdtd_dev_mode_overlap(fdtd_dev *d, int cx, int cy, int cz, const double *mode, int mi, int mj, double mdisc,
                          int reverse, double out[2]) {
    if (cz < 0 || cz >= d->K) {
        set_error("overlap plane outside the grid");
        return -2;
    }
    int err;
    const double *dm = mode_on_device(d, mode, mi, mj, mdisc, &err);
    if (!dm) return err;
    const unsigned long long ver = mode_cached(d, mode, mi, mj, mdisc)->version;
    const OvlReq req{cx, cy, cz, mi, mj, reverse, mdisc, mode};
    auto same = [](const OvlReq &x, const OvlReq &y) {
        return x.cx == y.cx && x.cy == y.cy && x.cz == y.cz && x.mi == y.mi && x.mj == y.mj &&
               x.reverse == y.reverse && x.mdisc == y.mdisc && x.host_key == y.host_key;
    };
    // the requests since the fields last changed form the pattern of the next batch
    if (d->seen_epoch != d->field_epoch) {
        if (d->n_seen > 0) {
            for (int q = 0; q < d->n_seen; q++) d->ovl_prev[q] = d->ovl_seen[q];
            d->n_prev = d->n_seen;
        }
        d->n_seen = 0;
        d->seen_epoch = d->field_epoch;
    }
    bool seen = false;
    for (int q = 0; q < d->n_seen; q++) seen = seen || same(d->ovl_seen[q], req);
    if (!seen && d->n_seen < kOvlSlots) d->ovl_seen[d->n_seen++] = req;
    // evaluated by the batch launched since the fields last changed, on the same mode content
    if (d->batch_epoch == d->field_epoch)
        for (int q = 0; q < d->n_batch; q++)
            if (same(d->ovl_req[q], req) && d->ovl_ver[q] == ver) {
                const volatile double2 *r = d->ovl_host + q;
                out[0] = r->x;
                out[1] = r->y;
                return 0;
            }
    // a new batch: this request, then the other requests of the last pattern whose modes are
    // still cached (their results wait for the calls that ask for them)
    OvlReq batch[kOvlSlots];
    const double *bmode[kOvlSlots];
    unsigned long long bver[kOvlSlots];
    int n = 0;
    batch[n] = req, bmode[n] = dm, bver[n++] = ver;
    for (int q = 0; q < d->n_prev && n < kOvlSlots; q++) {
        const OvlReq &o = d->ovl_prev[q];
        if (same(o, req)) continue;
        const ModeCache *mc = mode_cached(d, o.host_key, o.mi, o.mj, o.mdisc);
        if (!mc) continue;
        batch[n] = o, bmode[n] = mc->dev, bver[n++] = mc->version;
    }
    const int nb = (int)((d->P + kBlock - 1) / kBlock);
    if ((long long)nb * kOvlSlots > d->partial_cap) {
        (void)dev_free(d->partials);
        d->partials = nullptr;
        FDTD_TRY(dev_malloc(&d->partials, (size_t)nb * kOvlSlots * sizeof(double2)));
        d->partial_cap = nb * kOvlSlots;
    }
    // one launch per overlap, each writing straight into its slot of pinned host memory, and one
    // synchronisation for the batch: the reference returns each overlap from its own call
    for (int q = 0; q < n; q++) {
        const OvlReq &o = batch[q];
        hipLaunchKernelGGL(k_overlap, dim3(nb), dim3(kBlock), 0, d->stream, d->F[EX], d->F[EY], d->F[HX], d->F[HY],
                           d->I, d->J, d->P, o.cz, o.cx, o.cy, bmode[q], o.mi, o.mj, d->disc / o.mdisc, o.reverse,
                           d->partials + (size_t)q * nb, d->ovl_count + q, -0.5 * d->disc * d->disc, d->ovl_dev + q);
        FDTD_TRY(hipGetLastError());
    }
    FDTD_TRY(hipStreamSynchronize(d->stream));
    for (int q = 0; q < n; q++) d->ovl_req[q] = batch[q], d->ovl_ver[q] = bver[q];
    d->n_batch = n;
    d->batch_epoch = d->field_epoch;
    const volatile double2 *r = d->ovl_host;
    out[0] = r->x;
    out[1] = r->y;
    return 0;
}

int fdtd_dev_mode_overlap_noalloc(fdtd_dev *d, int cx, int cy, int cz, const double *mode_dev, int mi, int mj,
                                  double mdisc, int reverse, double *out_dev) {
    if (!d || !mode_dev || !out_dev) {
        set_error("overlap (no-alloc): null grid, mode or output pointer");
        return -2;
    }
    if (reinterpret_cast<uintptr_t>(out_dev) % alignof(double2) != 0) {
        set_error("overlap (no-alloc): the output must be 16-byte aligned device memory");
        return -2;
    }
    if (cz < 0 || cz >= d->K || mi <= 0 || mj <= 0 || !(mdisc > 0)) {
        set_error("overlap (no-alloc): plane outside the grid or empty mode window");
        return -2;
    }
    const long long nb = (d->P + kBlock - 1) / kBlock;
    if (nb > kNoAllocPartials) {
        set_error("overlap (no-alloc): planes are limited to " + std::to_string((long long)kNoAllocPartials * kBlock) +
                  " cells (this grid has " + std::to_string(d->P) + ")");
        return -2;
    }
    double2 *s = nullptr;
    const int rc = noalloc_scratch(d->stream, &s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_overlap, dim3((unsigned)nb), dim3(kBlock), 0, d->stream, d->F[EX], d->F[EY], d->F[HX],
                       d->F[HY], d->I, d->J, d->P, cz, cx, cy, mode_dev, mi, mj, d->disc / mdisc, reverse, s,
                       reinterpret_cast<unsigned *>(s + kNoAllocPartials), -0.5 * d->disc * d->disc,
                       reinterpret_cast<double2 *>(out_dev));
    FDTD_TRY(hipGetLastError());
    return 0;
}

int fdtd_dev_dft(fdtd_dev *d, double tw_re, double tw_im, int y_slice, int k0, int k1) {
    if (!d->fourier) {
        FDTD_TRY(dev_malloc(&d->fourier, (size_t)6 * d->N * sizeof(double2)));
        FDTD_TRY(hipMemsetAsync(d->fourier, 0, (size_t)6 * d->N * sizeof(double2), d->stream));
    }
    if (y_slice >= d->J) {
        set_error("y slice outside the grid");
        return -2;
    }
    k0 = k0 < 0 ? 0 : k0;
    k1 = k1 > d->K ? d->K : k1;
    if (k1 <= k0) return 0;
    long long n = y_slice >= 0 ? (long long)d->I * (k1 - k0) : d->P * (k1 - k0);
    unsigned nb = (unsigned)((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_dft, dim3(nb), dim3(kBlock), 0, d->stream, d->fourier, d->field_base, d->Nal, d->N, d->I,
                       d->J, d->P, k0, k1, y_slice, tw_re, tw_im);
    FDTD_TRY(hipGetLastError());
    return 0;
}

int fdtd_dev_taper_store(fdtd_dev *d, const void *key, int which, int seg_num, int start, int end, const int *left,
                         const int *right, int bottom, int top, int duration, int timestep) {
    const int L = end - start, H = top - bottom;
    if (L <= 0 || H <= 0 || duration <= 0 || seg_num <= 0) {
        set_error("empty taper");
        return -2;
    }
    if (timestep < 0 || timestep >= duration) {
        set_error("taper timestep outside [0, duration)");
        return -2;
    }
    auto it = g_tapers.find(key);
    if (it != g_tapers.end() && (it->second.H != H || it->second.L != L || it->second.T != duration)) {
        fdtd_dev_taper_release(key);
        it = g_tapers.end();
    }
    if (it == g_tapers.end()) {
        TaperHist th;
        th.key = key;
        th.H = H;
        th.L = L;
        th.T = duration;
        size_t n = (size_t)12 * H * L * duration;
        FDTD_TRY(dev_malloc(&th.hist[0], n * sizeof(float)));
        FDTD_TRY(dev_malloc(&th.hist[1], n * sizeof(float)));
        FDTD_TRY(hipMemset(th.hist[0], 0, n * sizeof(float)));
        FDTD_TRY(hipMemset(th.hist[1], 0, n * sizeof(float)));
        FDTD_TRY(dev_malloc(&th.edge_i, (size_t)2 * L * sizeof(int)));
        it = g_tapers.emplace(key, th).first;
    }
    TaperHist &th = it->second;
    // edge pixel per row: e[n] + round(int((e[n+1]-e[n])*(k-s0)/(s1-s0))) (adjoint3d.c:276,285)
    std::vector<int> edge((size_t)2 * L);
    for (int n = 0; n < seg_num; n++) {
        int s0 = start + (n * L) / seg_num, s1 = start + ((n + 1) * L) / seg_num;
        for (int k = s0; k < s1; k++) {
            for (int side = 0; side < 2; side++) {
                const int *e = side == 0 ? right : left;
                int i = (int)(e[n] + std::round((double)(((e[n + 1] - e[n]) * (k - s0)) / (s1 - s0))));
                if (i < 0 || i >= d->I) {
                    set_error("taper edge pixel outside the grid");
                    return -2;
                }
                edge[(size_t)side * L + (k - start)] = i;
            }
        }
    }
    if (edge != th.edge_host) {
        th.edge_host = edge;
        FDTD_TRY(hipMemcpyAsync(th.edge_i, th.edge_host.data(), edge.size() * sizeof(int), hipMemcpyHostToDevice,
                                d->stream));
    }
    if (bottom < 0 || top > d->J || start < 0 || end > d->K) {
        set_error("taper box outside the grid");
        return -2;
    }
    unsigned nb = (unsigned)((2 * H * L + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_taper_gather, dim3(nb), dim3(kBlock), 0, d->stream, th.hist[which ? 1 : 0], d->field_base,
                       d->Nal, th.edge_i, d->I, d->J, start, L, bottom, H, timestep);
    FDTD_TRY(hipGetLastError());
    return 0;
}

int fdtd_dev_taper_ep_gradient(const void *key, int height, int length, int duration, double *out) {
    auto it = g_tapers.find(key);
    if (it == g_tapers.end()) {
        std::memset(out, 0, sizeof(double) * 2 * height * length);
        return 0;
    }
    TaperHist &th = it->second;
    if (th.H != height || th.L != length || th.T != duration) {
        set_error("taper dimensions changed since the fields were stored");
        return -2;
    }
    double *g = nullptr;
    FDTD_TRY(hipDeviceSynchronize());
    FDTD_TRY(dev_malloc(&g, sizeof(double) * 2 * height * length));
    unsigned nb = (unsigned)((2 * height * length + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_taper_grad, dim3(nb), dim3(kBlock), 0, 0, th.hist[0], th.hist[1], height, length, duration,
                       g);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(out, g, sizeof(double) * 2 * height * length, hipMemcpyDeviceToHost);
    (void)dev_free(g);
    if (e != hipSuccess) return fail("fdtd_dev_taper_ep_gradient", e);
    return 0;
}

void fdtd_dev_taper_release(const void *key) {
    auto it = g_tapers.find(key);
    if (it == g_tapers.end()) return;
    (void)hipDeviceSynchronize();
    (void)dev_free(it->second.hist[0]);
    (void)dev_free(it->second.hist[1]);
    (void)dev_free(it->second.edge_i);
    g_tapers.erase(it);
}

int fdtd_dev_block_store(fdtd_dev *d, const void *key, const int box[6], int duration, int timestep) {
    const int nx = box[1] - box[0], ny = box[3] - box[2], nz = box[5] - box[4];
    if (nx <= 0 || ny <= 0 || nz <= 0 || box[0] < 0 || box[1] > d->I || box[2] < 0 || box[3] > d->J ||
        box[4] < 0 || box[5] > d->K) {
        set_error("adjoint block outside the grid");
        return -2;
    }
    if (timestep < 0 || timestep >= duration) {
        set_error("block timestep outside [0, t)");
        return -2;
    }
    auto it = g_blocks.find(key);
    if (it == g_blocks.end()) {
        BlockHist bh;
        bh.key = key;
        std::memcpy(bh.box, box, sizeof bh.box);
        bh.T = duration;
        size_t n = (size_t)nx * ny * nz;
        FDTD_TRY(dev_malloc(&bh.fwd, 3 * n * duration * sizeof(float)));
        FDTD_TRY(hipMemset(bh.fwd, 0, 3 * n * duration * sizeof(float)));
        FDTD_TRY(dev_malloc(&bh.grad, n * sizeof(double)));
        FDTD_TRY(hipMemset(bh.grad, 0, n * sizeof(double)));
        it = g_blocks.emplace(key, bh).first;
    }
    long long n = (long long)nx * ny * nz;
    unsigned nb = (unsigned)((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_block_gather, dim3(nb), dim3(kBlock), 0, d->stream, it->second.fwd, d->field_base, d->Nal,
                       d->I, d->J, box[0], nx, box[2], ny, box[4], nz, timestep);
    FDTD_TRY(hipGetLastError());
    return 0;
}

int fdtd_dev_block_update_gradient(fdtd_dev *d, const void *key, const int box[6], int duration, int timestep,
                                   double *gradient_host) {
    auto it = g_blocks.find(key);
    const int nx = box[1] - box[0], ny = box[3] - box[2], nz = box[5] - box[4];
    long long n = (long long)nx * ny * nz;
    if (it == g_blocks.end()) {
        // no stored forward fields: the reference multiplies by zeros
        return 0;
    }
    if (timestep < 0 || timestep >= duration) {
        set_error("block timestep outside [0, t)");
        return -2;
    }
    unsigned nb = (unsigned)((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_block_accumulate, dim3(nb), dim3(kBlock), 0, d->stream, it->second.grad, it->second.fwd,
                       d->field_base, d->Nal, d->I, d->J, box[0], nx, box[2], ny, box[4], nz, timestep);
    FDTD_TRY(hipGetLastError());
    if (gradient_host) {
        FDTD_TRY(hipMemcpyAsync(gradient_host, it->second.grad, n * sizeof(double), hipMemcpyDeviceToHost,
                                d->stream));
        FDTD_TRY(hipStreamSynchronize(d->stream));
    }
    return 0;
}

int fdtd_dev_block_reset(const void *key, int which) {
    auto it = g_blocks.find(key);
    if (it == g_blocks.end()) return 0;
    const int *box = it->second.box;
    size_t n = (size_t)(box[1] - box[0]) * (box[3] - box[2]) * (box[5] - box[4]);
    FDTD_TRY(hipDeviceSynchronize());
    if (which == 0) FDTD_TRY(hipMemset(it->second.grad, 0, n * sizeof(double)));
    else FDTD_TRY(hipMemset(it->second.fwd, 0, 3 * n * it->second.T * sizeof(float)));
    return 0;
}

void fdtd_dev_block_release(const void *key) {
    auto it = g_blocks.find(key);
    if (it == g_blocks.end()) return;
    (void)hipDeviceSynchronize();
    (void)dev_free(it->second.fwd);
    (void)dev_free(it->second.grad);
    g_blocks.erase(it);
}

}  // extern "C"
