// Device-resident runs: tabulated mode sources, FP64 overlap monitors with a multi-frequency
// running DFT, FP64 multi-frequency DFT planes.  Extension of the reference's per-call API
// (timefield3d.c:2977-3239): the same source and overlap formulas, but the per-step amplitude,
// phase and twiddles come from tables, so a run advances many steps without host round trips.
#include <cmath>
#include <cstring>
#include <vector>

#include "dev_leap.h"
#include "dev_stencil.h"

namespace fdtd {
namespace {

struct RunSource {
    int k;
    float *PR, *PI;        // 6 x P profile planes: J_m(n) = c1(n) PR_m + c2(n) PI_m
    unsigned char *win;    // P, 1 inside the mode window
    std::vector<float2> coef;  // nsteps x (A cos(phase), A sin(phase)), passed as kernel arguments
    int nsteps;
    float2 *dcoef = nullptr;   // device copy of coef for batched (step-counter) launches
};

struct RunOverlap {
    int k;
    float *W;              // 8 x P: Re/Im of mEx, mEy, mHx, mHy (reverse applied), 0 outside window
    int nfreq;
    std::vector<double> omega;
    double2 *acc;          // nfreq cumulative overlaps
    double2 *inst;         // last instantaneous overlap
    int freq_off;          // offset into the run's twiddle table
    bool ring_free;        // plane inside [1, K-2] and window clear of the i/j boundary ring
};

// A DFT monitor over the box [i0,i1) x [j0,j1) x [k0,k1) (a plane is the box of one z layer).
struct RunDft {
    int box[6];
    long long nb;          // cells in the box
    int mask, ncomp, nfreq;
    std::vector<double> omega;
    double2 *acc;          // ncomp x nfreq x nb, box cells x fastest
    int freq_off;
};

// Per-monitor device descriptor for the merged overlap kernels.
struct MonDev {
    long long koff;   // plane offset k * P
    const float *W;   // 8 x P weights
    double2 *inst;
    double2 *acc;
    int nfreq;
    int twoff;        // offset of this monitor's frequencies in a twiddle row
};

struct Run {
    std::vector<RunSource> src;
    std::vector<RunOverlap> ovl;
    std::vector<RunDft> dft;
    double2 *partials = nullptr;  // overlap stage-1 partial sums, [monitor][block]
    int partial_cap = 0;
    MonDev *mons = nullptr;       // device copy of the monitor descriptors
    int mon_cap = 0;
    std::vector<MonDev> mons_host;  // what `mons` holds (batched calls upload only a change)
    double2 *tw = nullptr;        // [step][total freq] twiddles for the current run() call
    long long tw_cap = 0;
    // monitors evaluated inside the H update: per-step block partials [monitor][step][block]
    // and per-step overlaps [monitor][step] of the current run() call
    double2 *mpart = nullptr;
    double2 *minst = nullptr;
    long long mstep_cap = 0;
    int mslot_cap = 0;  // partial slots per step that mpart holds (plane blocks or leap tiles)
};

// Monitor partial buffers for n steps of `slots` partials per monitor and step (layout
// [monitor][step][slot], step stride slots); grows, after synchronising st, when too small.
int ensure_monitor_parts(Run *r, int n, int slots, hipStream_t st) {
    if (n <= r->mstep_cap && slots <= r->mslot_cap) return 0;
    FDTD_TRY(hipStreamSynchronize(st));
    const long long cap = std::max<long long>(n, r->mstep_cap);
    const int sc = std::max(slots, r->mslot_cap);
    (void)dev_free(r->mpart);
    (void)dev_free(r->minst);
    r->mpart = r->minst = nullptr;
    r->mstep_cap = 0;
    r->mslot_cap = 0;
    FDTD_TRY(dev_malloc(&r->mpart, (size_t)kMaxRunMon * cap * sc * sizeof(double2)));
    FDTD_TRY(dev_malloc(&r->minst, (size_t)kMaxRunMon * cap * sizeof(double2)));
    r->mstep_cap = cap;
    r->mslot_cap = sc;
    return 0;
}

Run *run_of(fdtd_dev *d) {
    if (!d->run) d->run = new Run();
    return static_cast<Run *>(d->run);
}

__global__ void k_run_source(float *J, long long cs, const float *PR, const float *PI, const unsigned char *win,
                             long long P, float2 c) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P || !win[p]) return;
#pragma unroll
    for (int m = 0; m < 6; m++) J[m * cs + p] = c.x * PR[m * P + p] + c.y * PI[m * P + p];
}

// The plane of a run source for a leap launch: J (components 0-2) at the E half's step, M
// (components 3-5) at the H half's step, inside the mode window.
__global__ void k_run_source_halves(float *J, long long cs, const float *PR, const float *PI, const unsigned char *win,
                                    long long P, float2 cJ, float2 cM) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P || !win[p]) return;
#pragma unroll
    for (int m = 0; m < 6; m++) {
        const float2 c = m < 3 ? cJ : cM;
        J[m * cs + p] = c.x * PR[m * P + p] + c.y * PI[m * P + p];
    }
}

// All overlap monitors in one launch, blockIdx.y selects the monitor: per-block FP64 partial of
// Hx mEy - Hy mEx + Ey mHx - Ex mHy over the monitor plane (timefield3d.c:3195-3224).
__global__ __launch_bounds__(kBlock) void k_run_overlap_partial_multi(const float *base, long long Nal, long long P,
                                                                      const MonDev *mons, double2 *partials,
                                                                      int nbp) {
    __shared__ double sre[kBlock / 64], sim[kBlock / 64];
    const MonDev M = mons[blockIdx.y];
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double re = 0.0, im = 0.0;
    if (p < P) {
        const long long c = p + M.koff;
        const float *W = M.W;
        const double ex = base[c], ey = base[Nal + c], hx = base[3 * Nal + c], hy = base[4 * Nal + c];
        const double eEx_r = W[0 * P + p], eEx_i = W[1 * P + p], eEy_r = W[2 * P + p], eEy_i = W[3 * P + p];
        const double hHx_r = W[4 * P + p], hHx_i = W[5 * P + p], hHy_r = W[6 * P + p], hHy_i = W[7 * P + p];
        re = (hx * eEy_r - hy * eEx_r) + (ey * hHx_r - ex * hHy_r);
        im = (hx * eEy_i - hy * eEx_i) + (ey * hHx_i - ex * hHy_i);
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
        partials[(long long)blockIdx.y * nbp + blockIdx.x] = make_double2(r, s);
    }
}

// One block per monitor: fixed-order sum of its partials, scale, DFT accumulation.
__global__ __launch_bounds__(kBlock) void k_run_overlap_final_multi(const double2 *partials, int nbp, double scale,
                                                                    const MonDev *mons, const double2 *tw_row) {
    __shared__ double sre[kBlock], sim[kBlock];
    const MonDev M = mons[blockIdx.x];
    const double2 *pp = partials + (long long)blockIdx.x * nbp;
    double r = 0, s = 0;
    for (int q = threadIdx.x; q < nbp; q += kBlock) {
        r += pp[q].x;
        s += pp[q].y;
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
    const double vr = scale * sre[0], vi = scale * sim[0];
    if (threadIdx.x == 0) *M.inst = make_double2(vr, vi);
    for (int f = threadIdx.x; f < M.nfreq; f += kBlock) {
        double2 t = tw_row[M.twoff + f], a = M.acc[f];
        a.x += t.x * vr - t.y * vi;
        a.y += t.x * vi + t.y * vr;
        M.acc[f] = a;
    }
}

// Overlap of one monitor at step s from the block partials the H update stored: inst[s] = scale *
// sum_b part[s][b], blocks summed in index order.
__device__ __forceinline__ void mon_step_sum(const double2 *part, double2 *inst, int s, int nbp, double scale) {
    const double2 *pp = part + (long long)s * nbp;
    double r = 0, q = 0;
    for (int b = 0; b < nbp; b++) {
        r += pp[b].x;
        q += pp[b].y;
    }
    inst[s] = make_double2(scale * r, scale * q);
}

// Overlap per step of every monitor m of a grid: part [m][s][b] and inst [m][s], step capacity cap.
__global__ void k_run_mon_steps(const double2 *part, double2 *inst, int n, int nbp, long long cap, double scale) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x, m = blockIdx.y;
    if (s >= n) return;
    mon_step_sum(part + (long long)m * cap * nbp, inst + (long long)m * cap, s, nbp, scale);
}

// Running DFT of a run call's per-step overlaps x[0..n) of monitor M, by one block: acc[f] +=
// sum_s tw[s][f] x[s].  Thread t sums a fixed contiguous range of steps, the block combines the
// partial sums in a fixed tree, so the result does not depend on scheduling.  Also leaves the
// last step's overlap in *M.inst.
__device__ __forceinline__ void mon_dft_block(const MonDev &M, const double2 *x, int n, const double2 *tw, int nf) {
    __shared__ double sre[kBlock], sim[kBlock];
    const int per = (n + kBlock - 1) / kBlock;
    const int s0 = threadIdx.x * per, s1 = min(n, s0 + per);
    for (int f = 0; f < M.nfreq; f++) {
        double r = 0, q = 0;
        for (int s = s0; s < s1; s++) {
            const double2 t = tw[(long long)s * nf + M.twoff + f], v = x[s];
            r += t.x * v.x - t.y * v.y;
            q += t.x * v.y + t.y * v.x;
        }
        sre[threadIdx.x] = r;
        sim[threadIdx.x] = q;
        __syncthreads();
        for (int w = kBlock / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) {
                sre[threadIdx.x] += sre[threadIdx.x + w];
                sim[threadIdx.x] += sim[threadIdx.x + w];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            double2 a = M.acc[f];
            M.acc[f] = make_double2(a.x + sre[0], a.y + sim[0]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *M.inst = x[n - 1];
}

// The running DFT of every monitor of a grid, one block per monitor (inst [m][s], capacity cap).
__global__ __launch_bounds__(kBlock) void k_run_mon_dft(const MonDev *mons, const double2 *inst, int n, long long cap,
                                                        const double2 *tw, int nf) {
    const MonDev M = mons[blockIdx.x];
    mon_dft_block(M, inst + (long long)blockIdx.x * cap, n, tw, nf);
}

// One overlap monitor of a batched run call: its block partials [s][b] and overlaps [s], its
// device descriptor, and its grid's twiddle rows [s][nf].
struct BatchMon {
    const double2 *part;
    double2 *inst;
    const MonDev *mon;
    const double2 *tw;
    int nbp, nf;
    double scale;
};

// The finish of a batched run call for every monitor of every grid in two launches: per-step
// overlaps (blockIdx.y = monitor), then the running DFTs (one block per monitor).  The same
// per-monitor arithmetic as k_run_mon_steps and k_run_mon_dft.
__global__ void k_batch_mon_steps(const BatchMon *bm, int n) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const BatchMon &e = bm[blockIdx.y];
    mon_step_sum(e.part, e.inst, s, e.nbp, e.scale);
}

__global__ __launch_bounds__(kBlock) void k_batch_mon_dft(const BatchMon *bm, int n) {
    const BatchMon &e = bm[blockIdx.x];
    const MonDev M = *e.mon;
    mon_dft_block(M, e.inst, n, e.tw, e.nf);
}

// acc[slot][f][q] += field_m(cell q of the box) * tw[f] for each selected component m (FP64).
__global__ void k_run_dft_box(double2 *acc, const float *base, long long Nal, int I, int J, int i0, int ni, int j0,
                              int nj, int k0, long long nb, int mask, int nfreq, const double2 *tw) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nb) return;
    const long long row = q / ni;
    const int i = i0 + (int)(q - row * ni);
    const int j = j0 + (int)(row % nj);
    const int k = k0 + (int)(row / nj);
    const long long c = i + (long long)I * (j + (long long)J * k);
    int slot = 0;
    for (int m = 0; m < 6; m++) {
        if (!(mask >> m & 1)) continue;
        const double v = base[m * Nal + c];
        for (int f = 0; f < nfreq; f++) {
            double2 *a = acc + ((long long)slot * nfreq + f) * nb + q;
            double2 t = tw[f], x = *a;
            x.x += v * t.x;
            x.y += v * t.y;
            *a = x;
        }
        slot++;
    }
}

// grad[q] = Re[w * sum over the E components both monitors hold of F[q] * A[q]] at frequency f.
__global__ void k_dft_gradient(double *grad, const double2 *F, int fmask, int fnf, const double2 *A, int amask,
                               int anf, int f, long long nb, double wr, double wi) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nb) return;
    double sr = 0, si = 0;
    int fs = 0, as = 0;
    for (int m = 0; m < 6; m++) {
        const bool inf = fmask >> m & 1, ina = amask >> m & 1;
        if (m < 3 && inf && ina) {
            const double2 x = F[((long long)fs * fnf + f) * nb + q], y = A[((long long)as * anf + f) * nb + q];
            sr += x.x * y.x - x.y * y.y;
            si += x.x * y.y + x.y * y.x;
        }
        fs += inf;
        as += ina;
    }
    grad[q] = wr * sr - wi * si;
}

int mode_window(fdtd_dev *d, int n, int c, double ratio, int mn) {
    double dist = (n - c) * ratio;
    return (int)dist + mn / 2;
}

}  // namespace

void run_release(fdtd_dev *d) {
    if (!d->run) return;
    Run *r = static_cast<Run *>(d->run);
    (void)hipStreamSynchronize(d->stream);
    for (auto &s : r->src) {
        (void)dev_free(s.PR);
        (void)dev_free(s.PI);
        (void)dev_free(s.win);
        (void)dev_free(s.dcoef);
    }
    for (auto &o : r->ovl) {
        (void)dev_free(o.W);
        (void)dev_free(o.acc);
        (void)dev_free(o.inst);
    }
    for (auto &q : r->dft) (void)dev_free(q.acc);
    (void)dev_free(r->partials);
    (void)dev_free(r->mons);
    (void)dev_free(r->tw);
    (void)dev_free(r->mpart);
    (void)dev_free(r->minst);
    delete r;
    d->run = nullptr;
}

}  // namespace fdtd

using namespace fdtd;

extern "C" {

int fdtd_dev_run_clear(fdtd_dev *d) {
    run_release(d);
    return 0;
}

int fdtd_dev_run_add_source(fdtd_dev *d, int cx, int cy, int cz, const double *mode, int mi, int mj, double mdisc,
                            int reverse, int adjoint, int nsteps, const double *amplitude, const double *phase) {
    if (cz < 0 || cz >= d->K || nsteps <= 0) {
        set_error("run source: plane outside the grid or empty schedule");
        return -2;
    }
    if (!d->have_mats) {
        set_error("materials were never uploaded");
        return -1;
    }
    Run *r = run_of(d);
    const long long P = d->P;
    const double ratio = d->disc / mdisc;
    const int msign = -1 + 2 * reverse;
    // ep/mu of the plane cells from the material table
    std::vector<uint32_t> ids((size_t)P);
    {
        std::vector<unsigned char> raw((size_t)P * d->mat_bytes);
        FDTD_TRY(hipStreamSynchronize(d->stream));
        FDTD_TRY(hipMemcpy(raw.data(), (const char *)d->mat + (size_t)cz * P * d->mat_bytes, raw.size(),
                           hipMemcpyDeviceToHost));
        for (long long p = 0; p < P; p++)
            ids[p] = d->mat_bytes == 1   ? raw[p]
                     : d->mat_bytes == 2 ? reinterpret_cast<const uint16_t *>(raw.data())[p]
                                         : reinterpret_cast<const uint32_t *>(raw.data())[p];
    }
    std::vector<float> PR((size_t)6 * P, 0.0f), PI((size_t)6 * P, 0.0f);
    std::vector<unsigned char> win((size_t)P, 0);
    for (int j = 0; j < d->J; j++)
        for (int i = 0; i < d->I; i++) {
            int mx = mode_window(d, i, cx, ratio, mi), my = mode_window(d, j, cy, ratio, mj);
            if (mx < 0 || mx >= mi || my < 0 || my >= mj) continue;
            long long p = i + (long long)d->I * j;
            win[p] = 1;
            const double *md = mode + 12 * (mx + (long long)mi * my);
            const float *pr = &d->props[(size_t)ids[p] * kPropStride];
            if (!adjoint) {
                for (int m = 0; m < 6; m++) {
                    double f = m < 3 ? (double)pr[m] : (double)(-msign) * (double)pr[3 + m % 3];
                    PR[m * P + p] = (float)(f * md[2 * m]);
                    PI[m * P + p] = (float)(f * md[2 * m + 1]);
                }
            } else {
                const double fr[6][2] = {{md[8], md[9]},
                                         {-md[6], -md[7]},
                                         {0, 0},
                                         {(double)(-msign) * md[2], (double)(-msign) * md[3]},
                                         {(double)msign * md[0], (double)msign * md[1]},
                                         {0, 0}};
                for (int m = 0; m < 6; m++) {
                    PR[m * P + p] = (float)fr[m][0];
                    PI[m * P + p] = (float)fr[m][1];
                }
            }
        }
    std::vector<float2> coef((size_t)nsteps);
    for (int n = 0; n < nsteps; n++) {
        double A = adjoint ? 1.0 : amplitude[n];
        coef[n] = make_float2((float)(A * std::cos(phase[n])), (float)(A * std::sin(phase[n])));
    }
    int err;
    if (!plane_for(d, cz, &err)) return err;
    RunSource s;
    s.k = cz;
    s.nsteps = nsteps;
    s.coef = std::move(coef);
    FDTD_TRY(dev_malloc(&s.PR, PR.size() * sizeof(float)));
    FDTD_TRY(dev_malloc(&s.PI, PI.size() * sizeof(float)));
    FDTD_TRY(dev_malloc(&s.win, win.size()));
    FDTD_TRY(hipMemcpy(s.PR, PR.data(), PR.size() * sizeof(float), hipMemcpyHostToDevice));
    FDTD_TRY(hipMemcpy(s.PI, PI.data(), PI.size() * sizeof(float), hipMemcpyHostToDevice));
    FDTD_TRY(hipMemcpy(s.win, win.data(), win.size(), hipMemcpyHostToDevice));
    r->src.push_back(std::move(s));
    return (int)r->src.size() - 1;
}

int fdtd_dev_run_add_overlap(fdtd_dev *d, int cx, int cy, int cz, const double *mode, int mi, int mj, double mdisc,
                             int reverse, int nfreq, const double *omega) {
    if (cz < 0 || cz >= d->K || nfreq < 0) {
        set_error("run overlap: plane outside the grid");
        return -2;
    }
    Run *r = run_of(d);
    const long long P = d->P;
    const double ratio = d->disc / mdisc;
    std::vector<float> W((size_t)8 * P, 0.0f);
    for (int j = 0; j < d->J; j++)
        for (int i = 0; i < d->I; i++) {
            int mx = mode_window(d, i, cx, ratio, mi), my = mode_window(d, j, cy, ratio, mj);
            if (mx < 0 || mx >= mi || my < 0 || my >= mj) continue;
            long long p = i + (long long)d->I * j;
            const double *md = mode + 12 * (mx + (long long)mi * my);
            double v[8] = {md[0], md[1], md[2], md[3], md[6], md[7], md[8], md[9]};
            if (reverse == 1) {  // mH -> -conj(mH)
                v[4] = -v[4];
                v[6] = -v[6];
            }
            for (int q = 0; q < 8; q++) W[q * P + p] = (float)v[q];
        }
    RunOverlap o;
    o.k = cz;
    o.nfreq = nfreq;
    o.omega.assign(omega, omega + nfreq);
    o.freq_off = 0;
    // the H update only sees cells with 1 <= i <= I-2, 1 <= j <= J-2: the monitor can ride on it
    // when its weights vanish on the boundary ring and its plane is an updated one
    o.ring_free = cz >= 1 && cz <= d->K - 2;
    for (int j = 0; j < d->J && o.ring_free; j++)
        for (int i = 0; i < d->I; i++) {
            if (i != 0 && i != d->I - 1 && j != 0 && j != d->J - 1) continue;
            for (int q = 0; q < 8; q++)
                if (W[q * P + i + (long long)d->I * j] != 0.0f) o.ring_free = false;
        }
    const size_t nf1 = (size_t)(nfreq > 0 ? nfreq : 1);
    o.W = nullptr;
    o.acc = o.inst = nullptr;
    FDTD_TRY(dev_malloc(&o.W, W.size() * sizeof(float)));
    FDTD_TRY(hipMemcpy(o.W, W.data(), W.size() * sizeof(float), hipMemcpyHostToDevice));
    FDTD_TRY(dev_malloc(&o.acc, nf1 * sizeof(double2)));
    FDTD_TRY(hipMemset(o.acc, 0, nf1 * sizeof(double2)));
    FDTD_TRY(dev_malloc(&o.inst, sizeof(double2)));
    FDTD_TRY(hipMemset(o.inst, 0, sizeof(double2)));

    r->ovl.push_back(std::move(o));
    return (int)r->ovl.size() - 1;
}

int fdtd_dev_run_add_dft_box(fdtd_dev *d, const int box[6], int comp_mask, int nfreq, const double *omega) {
    const int N[3] = {d->I, d->J, d->K};
    for (int a = 0; a < 3; a++)
        if (box[2 * a] < 0 || box[2 * a + 1] > N[a] || box[2 * a] >= box[2 * a + 1]) {
            set_error("run DFT box: empty or outside the grid");
            return -2;
        }
    if (nfreq <= 0 || (comp_mask & 63) == 0) {
        set_error("run DFT box: bad component mask or frequency list");
        return -2;
    }
    Run *r = run_of(d);
    RunDft q;
    for (int a = 0; a < 6; a++) q.box[a] = box[a];
    q.nb = (long long)(box[1] - box[0]) * (box[3] - box[2]) * (box[5] - box[4]);
    q.mask = comp_mask & 63;
    q.ncomp = __builtin_popcount(q.mask);
    q.nfreq = nfreq;
    q.omega.assign(omega, omega + nfreq);
    q.freq_off = 0;
    q.acc = nullptr;
    size_t n = (size_t)q.ncomp * nfreq * q.nb;
    FDTD_TRY(dev_malloc(&q.acc, n * sizeof(double2)));
    FDTD_TRY(hipMemset(q.acc, 0, n * sizeof(double2)));
    r->dft.push_back(std::move(q));
    return (int)r->dft.size() - 1;
}

int fdtd_dev_run_add_dft_plane(fdtd_dev *d, int k, int comp_mask, int nfreq, const double *omega) {
    const int box[6] = {0, d->I, 0, d->J, k, k + 1};
    return fdtd_dev_run_add_dft_box(d, box, comp_mask, nfreq, omega);
}

// Exact adjoint source of an overlap monitor (same plane, mode window and reverse rule as
// fdtd_dev_run_add_overlap).  With the overlap's weights w (inst = w_E.E + w_H.H on the plane):
// J = w_E = dx^2/2 (mHy', -mHx', 0) and M = -e^{-i omega dt} w_H = e^{-i omega dt} dx^2/2 (mEy, -mEx, 0),
// mH' = -conj(mH) when reverse; step n writes Re[(J, M) amplitude[n] e^{-i phase[n]}].
int fdtd_dev_run_add_adjoint_source(fdtd_dev *d, int cx, int cy, int cz, const double *mode, int mi, int mj,
                                    double mdisc, int reverse, double omega_dt, int nsteps, const double *amplitude,
                                    const double *phase) {
    if (cz < 0 || cz >= d->K || nsteps <= 0) {
        set_error("run adjoint source: plane outside the grid or empty schedule");
        return -2;
    }
    Run *r = run_of(d);
    const long long P = d->P;
    const double ratio = d->disc / mdisc, h = 0.5 * d->disc * d->disc;
    const double zr = std::cos(omega_dt), zi = -std::sin(omega_dt);  // e^{-i omega dt}
    std::vector<float> PR((size_t)6 * P, 0.0f), PI((size_t)6 * P, 0.0f);
    std::vector<unsigned char> win((size_t)P, 0);
    for (int j = 0; j < d->J; j++)
        for (int i = 0; i < d->I; i++) {
            int mx = mode_window(d, i, cx, ratio, mi), my = mode_window(d, j, cy, ratio, mj);
            if (mx < 0 || mx >= mi || my < 0 || my >= mj) continue;
            long long p = i + (long long)d->I * j;
            win[p] = 1;
            const double *md = mode + 12 * (mx + (long long)mi * my);
            double hxr = md[6], hxi = md[7], hyr = md[8], hyi = md[9];
            if (reverse == 1) {  // mH -> -conj(mH)
                hxr = -hxr;
                hyr = -hyr;
            }
            const double Jv[3][2] = {{h * hyr, h * hyi}, {-h * hxr, -h * hxi}, {0, 0}};
            const double Mraw[3][2] = {{h * md[2], h * md[3]}, {-h * md[0], -h * md[1]}, {0, 0}};
            for (int m = 0; m < 3; m++) {
                PR[m * P + p] = (float)Jv[m][0];
                PI[m * P + p] = (float)Jv[m][1];
                PR[(3 + m) * P + p] = (float)(zr * Mraw[m][0] - zi * Mraw[m][1]);
                PI[(3 + m) * P + p] = (float)(zr * Mraw[m][1] + zi * Mraw[m][0]);
            }
        }
    std::vector<float2> coef((size_t)nsteps);
    for (int n = 0; n < nsteps; n++)
        coef[n] = make_float2((float)(amplitude[n] * std::cos(phase[n])), (float)(amplitude[n] * std::sin(phase[n])));
    int err;
    if (!plane_for(d, cz, &err)) return err;
    RunSource s;
    s.k = cz;
    s.nsteps = nsteps;
    s.coef = std::move(coef);
    s.PR = s.PI = nullptr;
    s.win = nullptr;
    FDTD_TRY(dev_malloc(&s.PR, PR.size() * sizeof(float)));
    FDTD_TRY(dev_malloc(&s.PI, PI.size() * sizeof(float)));
    FDTD_TRY(dev_malloc(&s.win, win.size()));
    FDTD_TRY(hipMemcpy(s.PR, PR.data(), PR.size() * sizeof(float), hipMemcpyHostToDevice));
    FDTD_TRY(hipMemcpy(s.PI, PI.data(), PI.size() * sizeof(float), hipMemcpyHostToDevice));
    FDTD_TRY(hipMemcpy(s.win, win.data(), win.size(), hipMemcpyHostToDevice));
    r->src.push_back(std::move(s));
    return (int)r->src.size() - 1;
}

}  // extern "C"

namespace fdtd {
namespace {

// Steps n0 .. n0+n-1 of a run (n at most kRunChunk, which bounds the twiddle table and the
// per-step monitor partials).
constexpr int kRunChunk = 4096;

// Steps n0..n0+n-1 as single-pass leap launches: E(n0) alone, then H(s) with E(s+1) for
// s = n0..n0+n-2 (reading one buffer, writing the other), then H(n0+n-1) alone.  The run sources
// are evaluated inside the launches with the coefficients of the step each half belongs to; the
// step-s overlap partials are formed by the launch that computes H(s).
int run_chunk_leap(fdtd_dev *d, Run *r, double dt, int n0, int n, int nf, const std::vector<SrcPlane *> &planes) {
    int rc = fused_prepare(d);
    if (rc) return rc;
    if ((rc = ensure_coefficients(d, dt)) || (rc = cpml_tables(d, dt))) return rc;
    fields_changed(d);
    const int kc = d->leap_kc > 0 ? d->leap_kc : leap_default_chunk(d);
    const int slots = leap_plane_slots(d);
    const int nmon = (int)r->ovl.size(), ns = (int)r->src.size();
    if (nmon > 0 && (rc = ensure_monitor_parts(r, n, slots, d->stream))) return rc;
    auto coef_at = [](const RunSource &rs, int step) {
        return step >= 0 && step < rs.nsteps ? rs.coef[step] : make_float2(0.0f, 0.0f);
    };
    const bool u8 = d->mat_bytes == 1;
    const unsigned nbp = (unsigned)((d->P + kBlock - 1) / kBlock);
    for (int j = 0; j <= n; j++) {
        const int mode = j == 0 ? 0 : (j == n ? 2 : 1);
        const int sh = n0 + j - 1;  // step of the H half (modes 1, 2)
        // each source plane holds J of the E half's step and M of the H half's step
        for (int q = 0; q < ns; q++) {
            const RunSource &rs = r->src[q];
            hipLaunchKernelGGL(k_run_source_halves, dim3(nbp), dim3(kBlock), 0, d->stream, planes[q]->J,
                               planes[q]->stride, rs.PR, rs.PI, rs.win, d->P, coef_at(rs, mode == 0 ? n0 : n0 + j),
                               coef_at(rs, sh));
            FDTD_TRY(hipGetLastError());
        }
        LeapArgs L;
        leap_fill(d, mode, kc, &L);
        L.nmon = mode == 0 ? 0 : nmon;
        for (int m = 0; m < L.nmon; m++)
            L.mon[m] = RunMonArg{r->ovl[m].k, r->ovl[m].W, r->mpart + ((size_t)m * r->mstep_cap + (j - 1)) * slots, 0};
        hipEvent_t t0 = nullptr, t1 = nullptr;
        if (mode == 1) timing_events(d, 2, &t0, &t1);
        if ((rc = leap_launch(L, u8, d->stream, t0, t1))) return rc;
        if (mode == 1) fused_swap(d);
        if (mode != 0 && (rc = finite_watch_scan(d, sh))) return rc;
    }
    const double scale = -0.5 * d->disc * d->disc;
    if (nmon > 0) {
        hipLaunchKernelGGL(k_run_mon_steps, dim3((n + kBlock - 1) / kBlock, nmon), dim3(kBlock), 0, d->stream,
                           (const double2 *)r->mpart, r->minst, n, slots, r->mstep_cap, scale);
        hipLaunchKernelGGL(k_run_mon_dft, dim3(nmon), dim3(kBlock), 0, d->stream, (const MonDev *)r->mons,
                           (const double2 *)r->minst, n, r->mstep_cap, (const double2 *)r->tw, nf);
        FDTD_TRY(hipGetLastError());
    }
    // leave each source's plane holding the values of the last step, as a per-step write would
    for (int q = 0; q < ns; q++) {
        const RunSource &rs = r->src[q];
        hipLaunchKernelGGL(k_run_source, dim3(nbp), dim3(kBlock), 0, d->stream, planes[q]->J, planes[q]->stride,
                           rs.PR, rs.PI, rs.win, d->P, coef_at(rs, n0 + n - 1));
        FDTD_TRY(hipGetLastError());
    }
    return 0;
}

int run_chunk(fdtd_dev *d, double dt, int n0, int n, double t_offset) {
    Run *r = run_of(d);
    // twiddle table for steps n0..n0+n-1 over every monitor frequency: e^{i omega (step + t_offset) dt}
    int nf = 0;
    for (auto &o : r->ovl) {
        o.freq_off = nf;
        nf += o.nfreq;
    }
    for (auto &q : r->dft) {
        q.freq_off = nf;
        nf += q.nfreq;
    }
    if (nf > 0) {
        std::vector<double2> tw((size_t)n * nf);
        for (int s = 0; s < n; s++) {
            double t = ((double)(n0 + s) + t_offset) * dt;
            int f = 0;
            for (auto &o : r->ovl)
                for (double w : o.omega) tw[(size_t)s * nf + f++] = make_double2(std::cos(w * t), std::sin(w * t));
            for (auto &q : r->dft)
                for (double w : q.omega) tw[(size_t)s * nf + f++] = make_double2(std::cos(w * t), std::sin(w * t));
        }
        if ((long long)tw.size() > r->tw_cap) {
            FDTD_TRY(hipStreamSynchronize(d->stream));
            (void)dev_free(r->tw);
            r->tw = nullptr;
            FDTD_TRY(dev_malloc(&r->tw, tw.size() * sizeof(double2)));
            r->tw_cap = (long long)tw.size();
        }
        FDTD_TRY(hipMemcpyAsync(r->tw, tw.data(), tw.size() * sizeof(double2), hipMemcpyHostToDevice, d->stream));
        FDTD_TRY(hipStreamSynchronize(d->stream));
    }
    const unsigned nbp = (unsigned)((d->P + kBlock - 1) / kBlock);
    const int nmon = (int)r->ovl.size();
    if (nmon > 0 && (long long)nbp * nmon > r->partial_cap) {
        FDTD_TRY(hipStreamSynchronize(d->stream));
        (void)dev_free(r->partials);
        r->partials = nullptr;
        FDTD_TRY(dev_malloc(&r->partials, (size_t)nbp * nmon * sizeof(double2)));
        r->partial_cap = (int)(nbp * nmon);
    }
    if (nmon > 0) {
        std::vector<MonDev> md((size_t)nmon);
        for (int m = 0; m < nmon; m++) {
            const RunOverlap &o = r->ovl[m];
            md[m] = MonDev{(long long)o.k * d->P, o.W, o.inst, o.acc, o.nfreq, o.freq_off};
        }
        if (nmon > r->mon_cap) {
            FDTD_TRY(hipStreamSynchronize(d->stream));
            (void)dev_free(r->mons);
            r->mons = nullptr;
            FDTD_TRY(dev_malloc(&r->mons, (size_t)nmon * sizeof(MonDev)));
            r->mon_cap = nmon;
        }
        FDTD_TRY(hipMemcpyAsync(r->mons, md.data(), md.size() * sizeof(MonDev), hipMemcpyHostToDevice, d->stream));
        FDTD_TRY(hipStreamSynchronize(d->stream));
        r->mons_host = md;
    }
    // source planes resolved once per call; the first kMaxRunSrc sources are evaluated inside
    // the Yee kernels, any further ones are written into their plane by a small kernel
    std::vector<SrcPlane *> planes;
    for (auto &s : r->src) {
        int err;
        SrcPlane *sp = plane_for(d, s.k, &err);
        if (!sp) return err;
        planes.push_back(sp);
    }
    auto coef_at = [](const RunSource &rs, int step) {
        // outside its schedule a source is switched off (amplitude 0)
        return step >= 0 && step < rs.nsteps ? rs.coef[step] : make_float2(0.0f, 0.0f);
    };
    const double scale = -0.5 * d->disc * d->disc;
    const int ninl = (int)r->src.size() < kMaxRunSrc ? (int)r->src.size() : kMaxRunSrc;
    if (d->fused_rows == -2 && leap_supported(d) && r->dft.empty() && (int)r->src.size() <= kMaxRunSrc &&
        nmon <= kMaxRunMon) {
        bool ring = true;
        for (auto &o : r->ovl) ring = ring && o.ring_free;
        if (ring) return run_chunk_leap(d, r, dt, n0, n, nf, planes);
    }
    // overlap monitors accumulate inside the H update when every one of them can
    bool fold = nmon > 0 && nmon <= kMaxRunMon && step_folds_monitors(d);
    for (auto &o : r->ovl) fold = fold && o.ring_free;
    const int mslots = step_monitor_slots(d);  // partials per step that the folding kernel writes
    if (fold) {
        int rc = ensure_monitor_parts(r, n, mslots, d->stream);
        if (rc) return rc;
    }
    for (int s = 0; s < n; s++) {
        const int step = n0 + s;
        d->n_run_src = ninl;
        for (int q = 0; q < ninl; q++) {
            const RunSource &rs = r->src[q];
            const float2 c = coef_at(rs, step);
            d->run_src[q] = RunSrcArg{rs.k, rs.PR, rs.PI, rs.win, c.x, c.y, nullptr, 0};
        }
        for (size_t q = ninl; q < r->src.size(); q++) {
            const RunSource &rs = r->src[q];
            hipLaunchKernelGGL(k_run_source, dim3(nbp), dim3(kBlock), 0, d->stream, planes[q]->J, planes[q]->stride,
                               rs.PR, rs.PI, rs.win, d->P, coef_at(rs, step));
            FDTD_TRY(hipGetLastError());
        }
        if (fold) {
            d->n_run_mon = nmon;
            for (int m = 0; m < nmon; m++)
                d->run_mon[m] = RunMonArg{r->ovl[m].k, r->ovl[m].W,
                                          r->mpart + ((size_t)m * r->mstep_cap + s) * mslots, 0};
        }
        int rc = fdtd_dev_step(d, dt, 0);
        d->n_run_src = 0;
        const bool consumed = d->n_run_mon == 0;
        d->n_run_mon = 0;
        if (rc) return rc;
        if (!consumed) {
            set_error("run monitors were not consumed by the H update");
            return -3;
        }
        if (nmon > 0 && !fold) {
            hipLaunchKernelGGL(k_run_overlap_partial_multi, dim3(nbp, nmon), dim3(kBlock), 0, d->stream,
                               (const float *)d->field_base, d->Nal, d->P, (const MonDev *)r->mons, r->partials,
                               (int)nbp);
            hipLaunchKernelGGL(k_run_overlap_final_multi, dim3(nmon), dim3(kBlock), 0, d->stream,
                               (const double2 *)r->partials, (int)nbp, scale, (const MonDev *)r->mons,
                               (const double2 *)(r->tw + (size_t)s * nf));
            FDTD_TRY(hipGetLastError());
        }
        for (auto &q : r->dft) {
            hipLaunchKernelGGL(k_run_dft_box, dim3((unsigned)((q.nb + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                               d->stream, q.acc, (const float *)d->field_base, d->Nal, d->I, d->J, q.box[0],
                               q.box[1] - q.box[0], q.box[2], q.box[3] - q.box[2], q.box[4], q.nb, q.mask, q.nfreq,
                               (const double2 *)(r->tw + (size_t)s * nf + q.freq_off));
            FDTD_TRY(hipGetLastError());
        }
        if (int frc = finite_watch_scan(d, step)) return frc;
    }
    if (fold) {
        hipLaunchKernelGGL(k_run_mon_steps, dim3((n + kBlock - 1) / kBlock, nmon), dim3(kBlock), 0, d->stream,
                           (const double2 *)r->mpart, r->minst, n, mslots, r->mstep_cap, scale);
        hipLaunchKernelGGL(k_run_mon_dft, dim3(nmon), dim3(kBlock), 0, d->stream, (const MonDev *)r->mons,
                           (const double2 *)r->minst, n, r->mstep_cap, (const double2 *)r->tw, nf);
        FDTD_TRY(hipGetLastError());
    }
    // leave each inline source's plane holding the values of the last step, as a per-step
    // write would have
    for (int q = 0; q < ninl; q++) {
        const RunSource &rs = r->src[q];
        hipLaunchKernelGGL(k_run_source, dim3(nbp), dim3(kBlock), 0, d->stream, planes[q]->J, planes[q]->stride,
                           rs.PR, rs.PI, rs.win, d->P, coef_at(rs, n0 + n - 1));
        FDTD_TRY(hipGetLastError());
    }
    return 0;
}

// Offsets of every monitor's frequencies in a twiddle row (overlaps first, then DFT boxes);
// returns the row length.
int run_freq_offsets(Run *r) {
    int nf = 0;
    for (auto &o : r->ovl) {
        o.freq_off = nf;
        nf += o.nfreq;
    }
    for (auto &q : r->dft) {
        q.freq_off = nf;
        nf += q.nfreq;
    }
    return nf;
}

// Twiddle rows e^{i omega (step + t_offset) dt} of steps n0..n0+n-1, nf per row, into tw.
void twiddle_rows(const Run *r, double dt, int n0, int n, double t_offset, int nf, double2 *tw) {
    for (int s = 0; s < n; s++) {
        double t = ((double)(n0 + s) + t_offset) * dt;
        int f = 0;
        for (auto &o : r->ovl)
            for (double w : o.omega) tw[(size_t)s * nf + f++] = make_double2(std::cos(w * t), std::sin(w * t));
        for (auto &q : r->dft)
            for (double w : q.omega) tw[(size_t)s * nf + f++] = make_double2(std::cos(w * t), std::sin(w * t));
    }
}

}  // namespace

int run_batch_prepare(fdtd_dev *d, double dt, int n0, int n, const int *step, YeeArgs *a, PmlArgs *pm) {
    Run *r = run_of(d);
    if (!d->cpml.on || !step_uses_split(d)) {
        set_error("batched runs need grids under the PML boundary with the split step kernels");
        return -2;
    }
    if (r->src.size() > (size_t)kMaxRunSrc || r->ovl.size() > (size_t)kMaxRunMon || !r->dft.empty()) {
        set_error("batched runs take at most 2 run sources and 4 overlap monitors per grid, and no DFT monitors");
        return -2;
    }
    for (auto &o : r->ovl)
        if (!o.ring_free) {
            set_error("batched runs need overlap monitors clear of the boundary ring");
            return -2;
        }
    // the twiddle rows themselves are formed by run_batch_finish, while the steps run
    (void)run_freq_offsets(r);
    int rc = 0;
    const unsigned nbp = (unsigned)((d->P + kBlock - 1) / kBlock);
    const int nmon = (int)r->ovl.size();
    if (nmon > 0) {
        std::vector<MonDev> md((size_t)nmon);
        for (int m = 0; m < nmon; m++) {
            const RunOverlap &o = r->ovl[m];
            md[m] = MonDev{(long long)o.k * d->P, o.W, o.inst, o.acc, o.nfreq, o.freq_off};
        }
        // MonDev has no padding, so equal bytes mean equal descriptors
        if (md.size() != r->mons_host.size() || std::memcmp(md.data(), r->mons_host.data(), md.size() * sizeof(MonDev))) {
            if (nmon > r->mon_cap) {
                (void)dev_free(r->mons);
                r->mons = nullptr;
                r->mons_host.clear();
                FDTD_TRY(dev_malloc(&r->mons, (size_t)nmon * sizeof(MonDev)));
                r->mon_cap = nmon;
            }
            FDTD_TRY(hipMemcpy(r->mons, md.data(), md.size() * sizeof(MonDev), hipMemcpyHostToDevice));
            r->mons_host = md;
        }
        if ((rc = ensure_monitor_parts(r, n, (int)nbp, d->stream))) return rc;
    }
    for (auto &s : r->src) {
        int err;
        if (!plane_for(d, s.k, &err)) return err;
        if (!s.dcoef) {
            FDTD_TRY(dev_malloc(&s.dcoef, s.coef.size() * sizeof(float2)));
            FDTD_TRY(hipMemcpy(s.dcoef, s.coef.data(), s.coef.size() * sizeof(float2), hipMemcpyHostToDevice));
        }
    }
    if ((rc = fill_step_args(d, dt, a, pm))) return rc;
    a->nrs = (int)r->src.size();
    for (int q = 0; q < a->nrs; q++) {
        const RunSource &s = r->src[q];
        a->rs[q] = RunSrcArg{s.k, s.PR, s.PI, s.win, 0.0f, 0.0f, s.dcoef, s.nsteps};
    }
    set_source_range(*a);
    a->nmon = nmon;
    for (int m = 0; m < nmon; m++)
        a->mon[m] = RunMonArg{r->ovl[m].k, r->ovl[m].W, r->mpart + (size_t)m * r->mstep_cap * nbp, (long long)nbp};
    a->step = step;
    a->step0 = n0;
    return 0;
}

static_assert(sizeof(MonDev) == 40, "MonDev is compared bytewise: no padding");

struct BatchStage {
    char *host = nullptr;  // pinned: one call's twiddle rows of every grid, then its monitor table
    size_t host_cap = 0;
    char *dev = nullptr;   // device copy of the same bytes
    size_t dev_cap = 0;
};

BatchStage *batch_stage_new() { return new BatchStage(); }

void batch_stage_free(BatchStage *S) {
    if (!S) return;
    if (S->host) (void)hipHostFree(S->host);
    (void)dev_free(S->dev);
    delete S;
}

int run_batch_finish(fdtd_dev *const *grids, int ng, double dt, int n0, int n, hipStream_t st, BatchStage *S) {
    // staging layout: [twiddle rows of grid 0][grid 1]...[BatchMon table]
    std::vector<size_t> off((size_t)ng + 1, 0);
    std::vector<int> nfs((size_t)ng, 0);
    int nmon_all = 0;
    for (int g = 0; g < ng; g++) {
        Run *r = run_of(grids[g]);
        for (auto &o : r->ovl) nfs[g] += o.nfreq;
        off[g + 1] = off[g] + (size_t)n * nfs[g] * sizeof(double2);
        nmon_all += (int)r->ovl.size();
    }
    const size_t tab_off = off[ng], need = tab_off + (size_t)nmon_all * sizeof(BatchMon);
    if (nmon_all > 0) {
        // the previous call's copies from the staging ended with that call's synchronisation
        if (need > S->host_cap) {
            if (S->host) (void)hipHostFree(S->host);
            S->host = nullptr;
            S->host_cap = 0;
            FDTD_TRY(hipHostMalloc((void **)&S->host, need));
            S->host_cap = need;
        }
        if (need > S->dev_cap) {
            FDTD_TRY(hipStreamSynchronize(st));  // this call's steps do not read it; earlier finishes did
            (void)dev_free(S->dev);
            S->dev = nullptr;
            S->dev_cap = 0;
            FDTD_TRY(dev_malloc(&S->dev, need));
            S->dev_cap = need;
        }
        BatchMon *tab = reinterpret_cast<BatchMon *>(S->host + tab_off);
        int q = 0;
        for (int g = 0; g < ng; g++) {
            fdtd_dev *d = grids[g];
            Run *r = run_of(d);
            if (r->ovl.empty()) continue;
            twiddle_rows(r, dt, n0, n, 0.0, nfs[g], reinterpret_cast<double2 *>(S->host + off[g]));
            const long long nbp = (d->P + kBlock - 1) / kBlock;
            for (int m = 0; m < (int)r->ovl.size(); m++)
                tab[q++] = BatchMon{r->mpart + (size_t)m * r->mstep_cap * nbp, r->minst + (size_t)m * r->mstep_cap,
                                    r->mons + m, reinterpret_cast<const double2 *>(S->dev + off[g]), (int)nbp, nfs[g],
                                    -0.5 * d->disc * d->disc};
        }
        FDTD_TRY(hipMemcpyAsync(S->dev, S->host, need, hipMemcpyHostToDevice, st));
        const BatchMon *dtab = reinterpret_cast<const BatchMon *>(S->dev + tab_off);
        hipLaunchKernelGGL(k_batch_mon_steps, dim3((n + kBlock - 1) / kBlock, q), dim3(kBlock), 0, st, dtab, n);
        hipLaunchKernelGGL(k_batch_mon_dft, dim3(q), dim3(kBlock), 0, st, dtab, n);
        FDTD_TRY(hipGetLastError());
    }
    for (int g = 0; g < ng; g++) {
        fdtd_dev *d = grids[g];
        Run *r = run_of(d);
        const unsigned nbp = (unsigned)((d->P + kBlock - 1) / kBlock);
        for (auto &s : r->src) {
            int err;
            SrcPlane *sp = plane_for(d, s.k, &err);
            if (!sp) return err;
            const int last = n0 + n - 1;
            const float2 c = last >= 0 && last < s.nsteps ? s.coef[last] : make_float2(0.0f, 0.0f);
            hipLaunchKernelGGL(k_run_source, dim3(nbp), dim3(kBlock), 0, st, sp->J, sp->stride, s.PR, s.PI, s.win,
                               d->P, c);
            FDTD_TRY(hipGetLastError());
        }
    }
    return 0;
}

namespace {
}  // namespace
}  // namespace fdtd

extern "C" {

int fdtd_dev_run(fdtd_dev *d, double dt, int n0, int n, double t_offset) {
    int rc = finite_watch_reset(d);
    if (rc) return rc;
    for (int s = 0; s < n; s += kRunChunk) {
        rc = run_chunk(d, dt, n0 + s, n - s < kRunChunk ? n - s : kRunChunk, t_offset);
        if (rc) return rc;
    }
    return finite_watch_verdict(d);
}

int fdtd_dev_run_overlap_result(fdtd_dev *d, int id, double *cumulative, double *last_inst) {
    Run *r = run_of(d);
    if (id < 0 || id >= (int)r->ovl.size()) {
        set_error("no such overlap monitor");
        return -2;
    }
    FDTD_TRY(hipStreamSynchronize(d->stream));
    const RunOverlap &o = r->ovl[id];
    if (cumulative && o.nfreq > 0)
        FDTD_TRY(hipMemcpy(cumulative, o.acc, (size_t)o.nfreq * sizeof(double2), hipMemcpyDeviceToHost));
    if (last_inst) FDTD_TRY(hipMemcpy(last_inst, o.inst, sizeof(double2), hipMemcpyDeviceToHost));
    return 0;
}

int fdtd_dev_run_dft_result(fdtd_dev *d, int id, double *out) {
    Run *r = run_of(d);
    if (id < 0 || id >= (int)r->dft.size()) {
        set_error("no such DFT monitor");
        return -2;
    }
    FDTD_TRY(hipStreamSynchronize(d->stream));
    const RunDft &q = r->dft[id];
    FDTD_TRY(hipMemcpy(out, q.acc, (size_t)q.ncomp * q.nfreq * q.nb * sizeof(double2), hipMemcpyDeviceToHost));
    return 0;
}

int fdtd_dev_dft_gradient(fdtd_dev *df, int id_f, fdtd_dev *da, int id_a, int freq, double w_re, double w_im,
                          double *grad) {
    Run *rf = run_of(df), *ra = run_of(da);
    if (id_f < 0 || id_f >= (int)rf->dft.size() || id_a < 0 || id_a >= (int)ra->dft.size()) {
        set_error("DFT gradient: no such DFT monitor");
        return -2;
    }
    const RunDft &F = rf->dft[id_f], &A = ra->dft[id_a];
    for (int a = 0; a < 6; a++)
        if (F.box[a] != A.box[a]) {
            set_error("DFT gradient: the two monitors cover different boxes");
            return -2;
        }
    if (freq < 0 || freq >= F.nfreq || freq >= A.nfreq || (F.mask & A.mask & 7) == 0) {
        set_error("DFT gradient: frequency index out of range or no common E component");
        return -2;
    }
    FDTD_TRY(hipStreamSynchronize(da->stream));
    double *g = nullptr;
    FDTD_TRY(dev_malloc(&g, (size_t)F.nb * sizeof(double)));
    hipLaunchKernelGGL(k_dft_gradient, dim3((unsigned)((F.nb + kBlock - 1) / kBlock)), dim3(kBlock), 0, df->stream, g,
                       (const double2 *)F.acc, F.mask, F.nfreq, (const double2 *)A.acc, A.mask, A.nfreq, freq, F.nb,
                       w_re, w_im);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(df->stream);
    if (e == hipSuccess) e = hipMemcpy(grad, g, (size_t)F.nb * sizeof(double), hipMemcpyDeviceToHost);
    (void)dev_free(g);
    if (e != hipSuccess) return fail("DFT gradient", e);
    return 0;
}

}  // extern "C"
