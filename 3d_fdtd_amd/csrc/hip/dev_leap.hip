// Single-pass leapfrog launches on gfx950: H of step s and E of step s + 1 in one z-march, so every
// field component is read once and written once per cell-update (48 B of FP32 traffic, against
// 72 B for a split E / H pair).  Reference semantics of the two updates: timefield3d.c:1679-1757
// (H, forward differences) and 1759-1837 (E, backward differences), semi-implicit conductivity;
// the CPML terms are the extension stated in oracle/fdtd_oracle.c.
//
// Order.  The split step is E(s) = f(E(s-1), H(s-1)), then H(s) = g(H(s-1), E(s)).  A run of n steps
// is issued as E(s0) [mode 0], then n - 1 launches of H(s), E(s+1) [mode 1], then H(s0+n-1)
// [mode 2]: the same updates in the same order, so the fields equal the split kernels' to FP32
// rounding, and the step-s monitors are formed inside the launch that computes H(s), from E(s)
// (read) and H(s) (computed).
//
// Mapping.  A block of 16 wavefronts covers an extended tile of 64 x-columns by 16 y-rows and
// marches a z chunk; wavefront w holds row y0 - 1 + w and lane l column x0 - 1 + l, one cell per
// thread.  Row 0 and column 0 are the halo whose new H the E update needs (j-1, i-1); row th + 1
// and column tw + 1 the halo whose E the H update needs (j+1, i+1).  The i +/- 1 neighbours are
// lane shuffles; the j +/- 1 neighbours cross wavefronts through two double-buffered LDS planes
// (Ex, Ez of E(k) for the H update; Hx, Hz of the new H for the E update), written before the
// plane's single barrier.  The halo H values are recomputed from the same old inputs with the
// same arithmetic as their owner's, so mode 1 reads one buffer and writes the other (fields and
// CPML memories ping-pong).  The first plane of a z chunk recomputes the new H of the plane below
// (its owner is the previous chunk).  Boundary cells (index 0 or N-1 on any axis) are never
// written and hold the same values in both buffers.
//
// Per plane, a thread's compute issues no vector-memory load: the coefficient table is in LDS, the
// CPML factors of its column and row are in registers and those of the plane are scalar loads,
// and its inputs for plane k + 2 are issued while plane k is computed (three register slots whose
// roles the unrolled loop rotates, since copying a register whose load is in flight would wait
// for it; the vector-memory counter retires in order).
#include <algorithm>

#include "dev_leap.h"
#include "dev_stencil.h"

namespace fdtd {
namespace {

using Off = uint32_t;  // every buffer of a leap launch is addressed with 32-bit offsets

// CPML factors b, a, 1/kappa along one axis for E and for H (0, 0, 1 outside the slabs, where the
// corrected difference d / kappa + (b psi + a d) reduces to d).
struct Fac {
    float bE, aE, kE, bH, aH, kH;
};

__device__ __forceinline__ Fac fac_at(const float *tab, int N, int n) {
    return Fac{tab[n], tab[N + n], tab[2 * N + n], tab[3 * N + n], tab[4 * N + n], tab[5 * N + n]};
}

// Inputs of one plane for one thread: E of the plane above (z+1 for the H update, then the next
// plane's own E), old H, material id, CPML memories (H: x0 x1 y0 y1 z0 z1, then E).
template <typename MatT>
struct In {
    float e[3];
    float h[3];
    MatT m;  // converted where it is used: a conversion here would wait for the load
    float ps[12];
};

template <typename MatT, bool MON, bool LC>
__global__ __launch_bounds__(kLeapBlockW * kLeapBlockH) void k_leap(LeapArgs L) {
    constexpr int TW = kLeapBlockW, TH = kLeapBlockH;
    extern __shared__ float smem[];
    const int cfl = LC ? (L.nmat * kLeapCoef + 63) / 64 * 64 : 0;
    float *lcoef = smem;
    float *lE = smem + cfl;           // [buf][Ex, Ez][row][lane]
    float *lH = lE + 2 * 2 * TH * TW;  // [buf][Hx, Hz][row][lane]
    if (LC)
        for (int q = threadIdx.x; q < L.nmat * kLeapCoef; q += TW * TH)
            lcoef[q] = L.coef[(q / kLeapCoef) * kCoefStride + q % kLeapCoef];
    const int cst = LC ? kLeapCoef : kCoefStride;
    const float *cg = LC ? lcoef : L.coef;

    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tile = (int)(xcd_logical_block(blockIdx.x, gridDim.x) - L.vb0);
    const int wy = tile % L.nwy, t2 = tile / L.nwy;
    const int wx = t2 % L.nwx, kc = t2 / L.nwx;
    const int x0 = 1 + L.tw * wx, y0 = 1 + L.th * wy;
    const int xe = min(x0 + L.tw, L.I - 1), ye = min(y0 + L.th, L.J - 1);  // owned [x0, xe) x [y0, ye)
    const int i = x0 - 1 + lane, j = y0 - 1 + w;
    const bool hreg = i < xe && j < ye;                  // new H formed here (own cells, i-1 / j-1 halo)
    const bool own = lane >= 1 && w >= 1 && hreg;
    const int kb = 1 + kc * L.kc_len, ke = min(kb + L.kc_len, L.K - 1);
    const bool doH = L.mode != 0, doE = L.mode != 2;
    const Off I = (Off)L.I, J = (Off)L.J, P = (Off)L.P, nal = (Off)L.nal;
    // every thread reads a valid cell: E up to column xe / row ye, H up to xe - 1 / ye - 1 (past the
    // tile's last cell a thread reads that cell again, with no extra HBM traffic; those values
    // feed no stored or consumed result)
    const int ih = min(i, xe - 1), jh = min(j, ye - 1);
    const Off pe = (Off)min(i, xe) + I * (Off)min(j, ye), ph = (Off)ih + I * (Off)jh;
    const int sx = slab_index(ih, L.I, L.nlo[0], L.nhi[0]);
    const int sy = slab_index(jh, L.J, L.nlo[1], L.nhi[1]);  // uniform over the wavefront
    const int sxc = sx >= 0 ? sx : 0;
    const Fac fx = fac_at(L.tab + L.toff[0], L.I, ih), fy = fac_at(L.tab + L.toff[1], L.J, jh);
    // CPML memory offsets (axis, field, component) and this thread's index in the x / y slabs
    auto pso = [&](int ax, int f, int q) { return (Off)L.poff[ax] + (Off)(2 * f + q) * (Off)L.pn[ax]; };
    auto qx = [&](int k) { return (Off)sxc + (Off)L.W[0] * ((Off)jh + J * (Off)k); };
    auto qy = [&](int k) { return (Off)ih + I * ((Off)sy + (Off)L.W[1] * (Off)k); };
    auto qz = [&](int sz) { return ph + P * (Off)sz; };

    // A tile that touches an x slab loads the x memories for all its lanes (from a clamped valid
    // index where a lane lies outside the slab: its b = a = 0 ignore the value), so no load sits
    // behind a per-lane branch (the join after one waits for every load in flight).
    const bool xslab = x0 - 1 < L.nlo[0] || (L.nhi[0] > 0 && xe >= L.I - 1 - L.nhi[0]);
    auto load = [&](int k, In<MatT> &in) {
        const Off ce = pe + P * (Off)(k + 1);
        in.e[0] = ld(L.src, ce), in.e[1] = ld(L.src, nal + ce), in.e[2] = ld(L.src, 2 * nal + ce);
        const Off c = ph + P * (Off)k;
        in.h[0] = ld(L.src, 3 * nal + c), in.h[1] = ld(L.src, 4 * nal + c), in.h[2] = ld(L.src, 5 * nal + c);
        in.m = ld(static_cast<const MatT *>(L.mat), c);
#pragma unroll
        for (int q = 0; q < 12; q++) in.ps[q] = 0.0f;
        if (xslab) {  // uniform over the block
            const Off q = qx(k);
            if (doH) in.ps[0] = ld(L.psrc, pso(0, 1, 0) + q), in.ps[1] = ld(L.psrc, pso(0, 1, 1) + q);
            if (doE) in.ps[6] = ld(L.psrc, pso(0, 0, 0) + q), in.ps[7] = ld(L.psrc, pso(0, 0, 1) + q);
        }
        if (sy >= 0) {  // uniform over the wavefront
            const Off q = qy(k);
            if (doH) in.ps[2] = ld(L.psrc, pso(1, 1, 0) + q), in.ps[3] = ld(L.psrc, pso(1, 1, 1) + q);
            if (doE) in.ps[8] = ld(L.psrc, pso(1, 0, 0) + q), in.ps[9] = ld(L.psrc, pso(1, 0, 1) + q);
        }
        const int sz = slab_index(k, L.K, L.nlo[2], L.nhi[2]);
        if (sz >= 0) {  // uniform over the block
            const Off q = qz(sz);
            if (doH) in.ps[4] = ld(L.psrc, pso(2, 1, 0) + q), in.ps[5] = ld(L.psrc, pso(2, 1, 1) + q);
            if (doE) in.ps[10] = ld(L.psrc, pso(2, 0, 0) + q), in.ps[11] = ld(L.psrc, pso(2, 0, 1) + q);
        }
    };
    // components m0..m0+2 of this cell's source on plane k (stored planes: J, then M)
    auto source3 = [&](int k, int m0, float v[3]) {
        if (k < L.src_k0 || k >= L.src_k1) return false;  // uniform: most planes stop here
        const float *Jp = nullptr;
        for (int s = 0; s < L.nsrc; s++)
            if (L.srck[s] == k) Jp = L.srcJ[s];
        if (!Jp) return false;
        for (int m = 0; m < 3; m++) v[m] = ld(Jp, (Off)(m0 + m) * P + ph);
        return true;
    };

    float ek[3];  // E(k) of this thread's cell
    {
        const Off c = pe + P * (Off)(kb - 1);
        ek[0] = ld(L.src, c), ek[1] = ld(L.src, nal + c), ek[2] = ld(L.src, 2 * nal + c);
    }
    In<MatT> SA, SB, SC;
    load(kb - 1, SA);
    if (kb < ke) load(kb, SB);
    lE[(0 * TH + w) * TW + lane] = ek[0];
    lE[(1 * TH + w) * TW + lane] = ek[2];
    __syncthreads();  // the coefficient table and the first E plane
    float hpx = 0.0f, hpy = 0.0f;  // new Hx, Hy of the plane below

    // plane k from `cur` (its inputs), issuing plane k + 2's loads into `pre`
    auto plane = [&](int k, const In<MatT> &cur, In<MatT> &pre) {
        const int buf = (k - kb + 1) & 1;
        if (k + 2 < ke) load(k + 2, pre);
        const bool body = k >= kb;  // k = kb - 1: only the new H of the plane below the chunk
        const int sz = slab_index(k, L.K, L.nlo[2], L.nhi[2]);
        const Fac fz = fac_at(L.tab + L.toff[2], L.K, k);  // uniform: scalar loads
        float *le = lE + buf * 2 * TH * TW, *lh = lH + buf * 2 * TH * TW;
        // new H of plane k (the old H where no update applies: walls, halo corner, mode 0)
        float hn[3] = {cur.h[0], cur.h[1], cur.h[2]};
        const float ey_ip = __shfl_down(ek[1], 1, 64), ez_ip = __shfl_down(ek[2], 1, 64);
        const int wn = w + 1 < TH ? w + 1 : w;
        const float ex_jp = le[wn * TW + lane], ez_jp = le[(TH + wn) * TW + lane];
        if (doH && hreg && i >= 1 && j >= 1 && k >= 1) {
            // H <- caH H - cbH curl E - cjH M, CPML-corrected differences
            float x1 = ez_jp - ek[2], x2 = cur.e[1] - ek[1];
            float y1 = cur.e[0] - ek[0], y2 = ez_ip - ek[2];
            float z1 = ey_ip - ek[1], z2 = ex_jp - ek[0];
            const float *ps = cur.ps;
            const float nx0 = fx.bH * ps[0] + fx.aH * y2, nx1 = fx.bH * ps[1] + fx.aH * z1;
            const float ny0 = fy.bH * ps[2] + fy.aH * z2, ny1 = fy.bH * ps[3] + fy.aH * x1;
            const float nz0 = fz.bH * ps[4] + fz.aH * x2, nz1 = fz.bH * ps[5] + fz.aH * y1;
            x1 = x1 * fy.kH + ny1, x2 = x2 * fz.kH + nz0;
            y1 = y1 * fz.kH + nz1, y2 = y2 * fx.kH + nx0;
            z1 = z1 * fx.kH + nx1, z2 = z2 * fy.kH + ny0;
            if (body && own) {
                if (sx >= 0) {
                    const Off q = qx(k);
                    st(L.pdst, pso(0, 1, 0) + q, nx0);
                    st(L.pdst, pso(0, 1, 1) + q, nx1);
                }
                if (sy >= 0) {
                    const Off q = qy(k);
                    st(L.pdst, pso(1, 1, 0) + q, ny0);
                    st(L.pdst, pso(1, 1, 1) + q, ny1);
                }
                if (sz >= 0) {
                    const Off q = qz(sz);
                    st(L.pdst, pso(2, 1, 0) + q, nz0);
                    st(L.pdst, pso(2, 1, 1) + q, nz1);
                }
            }
            const float *cf = cg + (int)cur.m * cst + 9;
            hn[0] = cf[0] * hn[0] - cf[3] * (x1 - x2);
            hn[1] = cf[1] * hn[1] - cf[4] * (y1 - y2);
            hn[2] = cf[2] * hn[2] - cf[5] * (z1 - z2);
            float sv[3];
            if (source3(k, 3, sv)) {
                hn[0] -= cf[6] * sv[0];
                hn[1] -= cf[7] * sv[1];
                hn[2] -= cf[8] * sv[2];
            }
        }
        // this plane's new H for the row above, the next plane's E for the row below; one barrier
        lh[w * TW + lane] = hn[0];
        lh[(TH + w) * TW + lane] = hn[2];
        float *ln = lE + (buf ^ 1) * 2 * TH * TW;
        ln[w * TW + lane] = cur.e[0];
        ln[(TH + w) * TW + lane] = cur.e[2];
        __syncthreads();
        const float hy_im = __shfl_up(hn[1], 1, 64), hz_im = __shfl_up(hn[2], 1, 64);
        const int wp = w > 0 ? w - 1 : 0;
        const float hx_jm = lh[wp * TW + lane], hz_jm = lh[(TH + wp) * TW + lane];
        if (body && own) {
            const Off c = ph + P * (Off)k;
            if (doE) {
                // E <- caE E + cbE curl H - cjE J, CPML-corrected differences
                float x1 = hn[2] - hz_jm, x2 = hn[1] - hpy;
                float y1 = hn[0] - hpx, y2 = hn[2] - hz_im;
                float z1 = hn[1] - hy_im, z2 = hn[0] - hx_jm;
                const float *ps = cur.ps + 6;
                const float nx0 = fx.bE * ps[0] + fx.aE * y2, nx1 = fx.bE * ps[1] + fx.aE * z1;
                const float ny0 = fy.bE * ps[2] + fy.aE * z2, ny1 = fy.bE * ps[3] + fy.aE * x1;
                const float nz0 = fz.bE * ps[4] + fz.aE * x2, nz1 = fz.bE * ps[5] + fz.aE * y1;
                x1 = x1 * fy.kE + ny1, x2 = x2 * fz.kE + nz0;
                y1 = y1 * fz.kE + nz1, y2 = y2 * fx.kE + nx0;
                z1 = z1 * fx.kE + nx1, z2 = z2 * fy.kE + ny0;
                if (sx >= 0) {
                    const Off q = qx(k);
                    st(L.pdst, pso(0, 0, 0) + q, nx0);
                    st(L.pdst, pso(0, 0, 1) + q, nx1);
                }
                if (sy >= 0) {
                    const Off q = qy(k);
                    st(L.pdst, pso(1, 0, 0) + q, ny0);
                    st(L.pdst, pso(1, 0, 1) + q, ny1);
                }
                if (sz >= 0) {
                    const Off q = qz(sz);
                    st(L.pdst, pso(2, 0, 0) + q, nz0);
                    st(L.pdst, pso(2, 0, 1) + q, nz1);
                }
                const float *cf = cg + (int)cur.m * cst;
                float ex = cf[0] * ek[0] + cf[3] * (x1 - x2);
                float ey = cf[1] * ek[1] + cf[4] * (y1 - y2);
                float ez = cf[2] * ek[2] + cf[5] * (z1 - z2);
                float sv[3];
                if (source3(k, 0, sv)) {
                    ex -= cf[6] * sv[0];
                    ey -= cf[7] * sv[1];
                    ez -= cf[8] * sv[2];
                }
                st(L.dst, c, ex);
                st(L.dst, nal + c, ey);
                st(L.dst, 2 * nal + c, ez);
            }
            if (doH) {
                st(L.dst, 3 * nal + c, hn[0]);
                st(L.dst, 4 * nal + c, hn[1]);
                st(L.dst, 5 * nal + c, hn[2]);
            }
        }
        // step-s overlap monitors on this plane: E(s) read, H(s) just formed; one slot per wavefront
        if (MON && doH && body) {
            for (int q = 0; q < L.nmon; q++) {
                const RunMonArg &M = L.mon[q];
                if (M.k != k) continue;  // uniform over the block
                double re = 0.0, im = 0.0;
                if (own) {
                    double wv[8];
                    for (int r8 = 0; r8 < 8; r8++) wv[r8] = ld(M.W, (Off)r8 * P + ph);
                    const double ex = ek[0], ey = ek[1], hx = hn[0], hy = hn[1];
                    re = (hx * wv[2] - hy * wv[0]) + (ey * wv[4] - ex * wv[6]);
                    im = (hx * wv[3] - hy * wv[1]) + (ey * wv[5] - ex * wv[7]);
                }
                for (int off = 32; off > 0; off >>= 1) {
                    re += __shfl_down(re, off, 64);
                    im += __shfl_down(im, off, 64);
                }
                if (lane == 0) {
                    double2 *part = L.step ? M.part + (long long)(*L.step - L.step0) * M.stride : M.part;
                    part[(wy + L.nwy * wx) * TH + w] = make_double2(re, im);
                }
            }
        }
        hpx = hn[0], hpy = hn[1];
        ek[0] = cur.e[0], ek[1] = cur.e[1], ek[2] = cur.e[2];
    };
    for (int k = kb - 1; k < ke; k += 3) {
        plane(k, SA, SC);
        if (k + 1 >= ke) break;
        plane(k + 1, SB, SA);
        if (k + 2 >= ke) break;
        plane(k + 2, SC, SB);
    }
}

}  // namespace

bool leap_supported(const fdtd_dev *d) {
    const Cpml &cp = d->cpml;
    const long long pmem = 4 * (cp.psi_n[0] + cp.psi_n[1] + cp.psi_n[2]);
    return cp.on && cp.psi_base && !d->dense_src && d->mat_bytes <= 2 && d->I >= 3 && d->J >= 3 && d->K >= 3 && 6 * d->Nal < (1LL << 30) &&
           pmem < (1LL << 30) && (long long)d->src.size() <= kMaxSrcPlanes;
}

int leap_plane_slots(const fdtd_dev *d) {
    const int nwx = (d->I - 2 + kLeapBlockW - 3) / (kLeapBlockW - 2);
    const int nwy = (d->J - 2 + kLeapBlockH - 3) / (kLeapBlockH - 2);
    return nwx * nwy * kLeapBlockH;
}

int leap_default_chunk(const fdtd_dev *d) {
    (void)d;
    return 8;  // measured on the strip waveguide and the 12.4 M-cell taper (4-16 within a few %)
}

void leap_fill(fdtd_dev *d, int mode, int kc_len, LeapArgs *L) {
    LeapArgs &l = *L;
    l = LeapArgs{};
    const bool pp = mode == 1;  // ping-pong: write the second buffer
    l.src = d->field_base;
    l.dst = pp ? d->alt_base : d->field_base;
    l.nal = (int)d->Nal;
    l.mat = d->mat;
    l.coef = d->d_coef;  // semi-implicit scheme
    l.nmat = d->nmat <= kLeapLdsMats ? d->nmat : 0;
    const Cpml &cp = d->cpml;
    l.tab = cp.tab;
    long long off = 0;
    for (int ax = 0; ax < 3; ax++) {
        l.toff[ax] = (int)cp.tab_off[ax];
        l.nlo[ax] = cp.nlo[ax];
        l.nhi[ax] = cp.nhi[ax];
        l.W[ax] = cp.W[ax];
        l.poff[ax] = (int)off;  // the layout of fdtd_dev_cpml_on: per axis, E q0, E q1, H q0, H q1
        l.pn[ax] = (int)cp.psi_n[ax];
        off += 4 * cp.psi_n[ax];
    }
    l.psrc = cp.psi_base;
    l.pdst = pp ? cp.psi_alt : cp.psi_base;
    l.I = d->I;
    l.J = d->J;
    l.K = d->K;
    l.P = (int)d->P;
    l.mode = mode;
    l.nwx = (d->I - 2 + kLeapBlockW - 3) / (kLeapBlockW - 2);
    l.tw = (d->I - 2 + l.nwx - 1) / l.nwx;
    l.nwy = (d->J - 2 + kLeapBlockH - 3) / (kLeapBlockH - 2);
    l.th = (d->J - 2 + l.nwy - 1) / l.nwy;
    l.kc_len = kc_len;
    l.nkc = (d->K - 2 + kc_len - 1) / kc_len;
    l.nsrc = 0;
    int k0 = d->K, k1 = 0;
    for (auto &sp : d->src) {
        l.srck[l.nsrc] = sp.k;
        l.srcJ[l.nsrc++] = sp.J;
        k0 = std::min(k0, sp.k);
        k1 = std::max(k1, sp.k + 1);
    }
    l.src_k0 = k0 < k1 ? k0 : 0;
    l.src_k1 = k0 < k1 ? k1 : 0;
    l.nmon = 0;
    l.step = nullptr;
    l.step0 = 0;
    l.vb0 = 0;
}

long long leap_blocks(const LeapArgs &L) { return (long long)L.nwx * L.nwy * L.nkc; }

size_t leap_lds_bytes(const LeapArgs &L) {
    const int cfl = L.nmat > 0 ? (L.nmat * kLeapCoef + 63) / 64 * 64 : 0;
    return (size_t)(cfl + 2 * 2 * 2 * kLeapBlockH * kLeapBlockW) * sizeof(float);
}

int leap_launch(const LeapArgs &L, bool u8, hipStream_t st, hipEvent_t t0, hipEvent_t t1) {
    const dim3 grid((unsigned)leap_blocks(L)), block(kLeapBlockW * kLeapBlockH);
    const size_t lds = leap_lds_bytes(L);
    const bool lc = L.nmat > 0, mon = L.nmon > 0;
#define FDTD_LEAP(MT, MV, LV) hipExtLaunchKernelGGL((k_leap<MT, MV, LV>), grid, block, lds, st, t0, t1, 0, L)
#define FDTD_LEAP_L(MT, MV)                  \
    do {                                     \
        if (lc) FDTD_LEAP(MT, MV, true);     \
        else FDTD_LEAP(MT, MV, false);       \
    } while (0)
    if (u8 && mon) FDTD_LEAP_L(uint8_t, true);
    else if (u8) FDTD_LEAP_L(uint8_t, false);
    else if (mon) FDTD_LEAP_L(uint16_t, true);
    else FDTD_LEAP_L(uint16_t, false);
#undef FDTD_LEAP_L
#undef FDTD_LEAP
    FDTD_TRY(hipGetLastError());
    return 0;
}

}  // namespace fdtd
