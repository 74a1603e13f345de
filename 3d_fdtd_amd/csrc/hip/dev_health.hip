// Non-finite field detection.  The reference has no failure detection (its loops run on through
// NaN / Inf); a device-resident run here can scan the six field arrays every N steps and fail
// the run call with the count of non-finite values and the first step that showed them.  One
// scan reads 24 B per cell, so a scan every 100 steps costs about 0.3 % of the run.
#include <cmath>

#include "dev_internal.h"

namespace fdtd {
namespace {

struct FiniteBuf {
    unsigned long long count;
    int first_step;
    int pad;
};

// blockIdx.y = component; each block counts the non-finite values of its grid-stride share and
// one lane per block adds the count (vector atomics) and lowers the first step.
__global__ __launch_bounds__(kBlock) void k_count_nonfinite(const float *base, long long Nal, long long N,
                                                            FiniteBuf *out, int step) {
    __shared__ unsigned int part[kBlock / 64];
    const float *f = base + (long long)blockIdx.y * Nal;
    unsigned int n = 0;
    for (long long c = (long long)blockIdx.x * kBlock + threadIdx.x; c < N; c += (long long)gridDim.x * kBlock)
        n += !isfinite(f[c]);
    for (int off = 32; off > 0; off >>= 1) n += __shfl_down(n, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < kBlock / 64; w++) t += part[w];
        if (t) {
            atomicAdd(&out->count, t);
            atomicMin(&out->first_step, step);
        }
    }
}

int ensure_buf(fdtd_dev *d) {
    if (!d->finite_buf) FDTD_TRY(dev_malloc(&d->finite_buf, sizeof(FiniteBuf)));
    return 0;
}

int reset_buf(fdtd_dev *d) {
    int rc = ensure_buf(d);
    if (rc) return rc;
    static const FiniteBuf z{0ULL, 0x7fffffff, 0};  // outlives the asynchronous copy
    FDTD_TRY(hipMemcpyAsync(d->finite_buf, &z, sizeof z, hipMemcpyHostToDevice, d->stream));
    return 0;
}

int scan(fdtd_dev *d, int step) {
    const long long blocks = std::min<long long>((d->N + kBlock - 1) / kBlock, 1024);
    hipLaunchKernelGGL(k_count_nonfinite, dim3((unsigned)blocks, 6), dim3(kBlock), 0, d->stream,
                       (const float *)d->field_base, d->Nal, d->N, static_cast<FiniteBuf *>(d->finite_buf), step);
    FDTD_TRY(hipGetLastError());
    return 0;
}

int read_buf(fdtd_dev *d, FiniteBuf *h) {
    FDTD_TRY(hipMemcpyAsync(h, d->finite_buf, sizeof *h, hipMemcpyDeviceToHost, d->stream));
    FDTD_TRY(hipStreamSynchronize(d->stream));
    return 0;
}

}  // namespace

int finite_watch_reset(fdtd_dev *d) { return d->finite_every > 0 ? reset_buf(d) : 0; }

int finite_watch_scan(fdtd_dev *d, int step) {
    if (d->finite_every <= 0 || (step + 1) % d->finite_every != 0) return 0;
    return scan(d, step);
}

int finite_watch_verdict(fdtd_dev *d) {
    if (d->finite_every <= 0) return 0;
    FiniteBuf h;
    int rc = read_buf(d, &h);
    if (rc) return rc;
    if (h.count) {
        set_error("non-finite field values: " + std::to_string(h.count) + " after step " +
                  std::to_string(h.first_step));
        return -4;
    }
    return 0;
}

}  // namespace fdtd

using namespace fdtd;

extern "C" {

int fdtd_dev_count_nonfinite(fdtd_dev *d, long long *count) {
    int rc = reset_buf(d);
    if (!rc) rc = scan(d, -1);
    FiniteBuf h;
    if (!rc) rc = read_buf(d, &h);
    if (rc) return rc;
    if (count) *count = (long long)h.count;
    return 0;
}

int fdtd_dev_set_finite_check(fdtd_dev *d, int every) {
    if (every < 0) {
        set_error("finite check interval must be >= 0");
        return -2;
    }
    d->finite_every = every;
    return 0;
}

}  // extern "C"
