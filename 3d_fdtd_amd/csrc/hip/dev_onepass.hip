// Single-pass full steps on gfx950: E and then H of one time step in one z-march, so that every
// field component is read once and written once per cell-update (48 B of FP32 field traffic,
// against 72 B for the split E / H pair).  Reference semantics of the two updates:
// timefield3d.c:1759-1837 (E, backward differences) and 1679-1757 (H, forward differences),
// semi-implicit conductivity, in the order of timefield3d_apply_timestep_omp_full
// (timefield3d.c:2787-2849: E over the whole grid, then H); the CPML terms are the extension whose
// formulas oracle/fdtd_oracle.c states.  The per-cell arithmetic is the split kernels'
// (dev_stencil.h), operation for operation.
//
// Mapping.  A block of four wavefronts covers an extended x-y tile of 64 columns by 4 rows and
// marches a z chunk upward; a thread holds one column.  At plane k the block forms the new E of
// the tile's cells, then the new H of plane k - 1 of its tw x th owned cells (tw < 64, th < 4: the
// interior is split into tiles of equal size, so no tile is a sliver), whose E and H it stores.
// The new H of a cell needs the new E of its i+1 and j+1 neighbours: i+1 is the next lane (a DPP
// lane shift), j+1 the thread 64 further on, through a three-slot LDS ring written before the
// plane's only barrier; the column and row after the owned ones are computed only to serve those
// reads (their owner is the next tile), and lanes past them repeat them.  The old H of the i-1 /
// j-1 neighbours that the E update needs are plain loads (served by L1 / L2: the neighbouring
// threads load the same lines).  Fields and CPML memories ping-pong (read one buffer, write the
// other), so no thread reads a value another block has already replaced; boundary cells (index 0
// or N-1 on any axis) are never written and hold the same values in both buffers.  The top plane
// of a chunk repeats the next chunk's first E plane.
//
// Slab classes.  The blocks of a step are split by the CPML slabs their cells may lie in (clear,
// x slab, y slab, both, z-end chunks), one launch per class, so that each kernel carries only the
// memories it needs and keeps its own register budget (clear blocks: 79 VGPRs).  Isotropic
// lossless media (caE = caH = 1, one cbE per material, one cbH per grid) take a one-float
// coefficient gather per cell instead of the 12-float record.
//
// Latency.  Chunks of 4 planes issue every load of the chunk before the first plane is computed
// (ids, fields, then the coefficient gathers that need the ids), so each plane waits only for its
// own loads.  Other chunk lengths march with three rotating register sets, the loads of planes
// k + 1 and k + 2 in flight while plane k is computed.  No loop body has a branch around a memory
// instruction: the vector-memory counter retires in order, and the compiler can then count every
// wait exactly instead of draining the loads at a control-flow join.  Stores of cells a thread
// does not own, and CPML memories of cells outside the slabs, go through buffer descriptors with
// an out-of-range offset (the hardware drops such a store and returns 0 for such a load).
//
// Measured on MI355X it stays slower than the split pair (the 200 x 100 x 620 grid: 286 against
// 180 us per step): it moves fewer bytes (about 62 B per cell-update with CPML memories, halo
// re-reads and partial lines, against 85 B) but at 2.5 TB/s instead of 5.8, bound by the latency
// of each block's plane-by-plane march rather than by HBM.  It is selectable and parity-tested.
#include <algorithm>

#include "dev_onepass.h"

namespace fdtd {
namespace {

using Off = uint32_t;
constexpr unsigned kOob = 0x80000000u;  // offset past every buffer (all of them are below 2 GiB)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p, unsigned bytes) {
    // raw buffer (stride 0), 32-bit data format: offsets at or past `bytes` are out of range
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float buf_ld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
// streaming store (the written buffer is read again only by the next step)
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff, (int)soff, 2);
}
// lane l receives lane l + 1's value (the last lane its own)
__device__ __forceinline__ float lane_next(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, v), __builtin_bit_cast(int, v),
                                                                 0x130 /* wave_shl:1 */, 0xf, 0xf, false));
}

// Inputs of one thread at one plane k: old E, old H, the old H neighbours the E curl needs, the
// packed update coefficients of the cell's material, the material ids of plane k (sources) and of
// plane k + 1 (the address of the next plane's coefficients, so that their gather is issued with
// the next plane's other loads instead of after them), and the CPML memories (0 outside the
// slabs): E (x q0 q1, y q0 q1, z q0 q1), then H.
template <typename MatT>
struct OpIn {
    float eo[3];
    float h[3];
    float nb[4];  // Hy(i-1), Hz(i-1), Hx(j-1), Hz(j-1)
    float4 c0, c1, c2;  // packed coefficients (general tables)
    float cb;           // cbE (isotropic lossless tables)
    int m;
    MatT mn;  // converted where it is used (a conversion here would wait for the load)
    float pe[6];
    float ph[6];
};

// The values of plane k - 1 that the H update of that plane needs one iteration later.
struct OpPrev {
    float en[3];  // new E
    float h[3];   // old H
    float ca[3], cb[3];
    int m;
    float ph[6];  // old H memories
};

struct Fac {
    float bE, aE, kE, bH, aH, kH;
};
__device__ __forceinline__ Fac fac_at(const float *t, int N, int n) {
    return Fac{t[n], t[N + n], t[2 * N + n], t[3 * N + n], t[4 * N + n], t[5 * N + n]};
}

__device__ __forceinline__ int slab_of(int n, int N, int nlo, int nhi) {
    if (n < nlo) return n;
    if (nhi > 0 && n >= N - 1 - nhi) return nlo + n - (N - 1 - nhi);
    return -1;
}

// Block-wide FP64 sum over the 256 threads (fixed order); every thread must call it.
__device__ __forceinline__ double2 sum256(double re, double im, double *red) {
    for (int off = 32; off > 0; off >>= 1) {
        re += __shfl_down(re, off, 64);
        im += __shfl_down(im, off, 64);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        red[2 * w] = re;
        red[2 * w + 1] = im;
    }
    __syncthreads();
    const double2 s = make_double2((red[0] + red[2]) + (red[4] + red[6]), (red[1] + red[3]) + (red[5] + red[7]));
    __syncthreads();
    return s;
}

// SL: the CPML slab axes the block's cells may lie in (bit 0 x: the tile touches an x slab, bit 1
// y, bit 2 z: the chunk touches a z slab); the loop carries the memories of those axes only.
// KCU > 0: chunks of KCU planes whose loads are all issued before the first plane is computed;
// 0: chunks of A.kc planes marched with a rotating two-plane prefetch.
template <typename MatT, int TW, bool NT, bool ISO, int SL, int KCU>
__device__ __forceinline__ void onepass_block(const OnePassArgs &A, int tile, int kcix) {
    constexpr int TH = kOpBlock / TW;
    constexpr int kOwnAux = NT ? 2 : 0;  // streaming hint on the loads of a cell's own old E
    __shared__ float ring[3][2][kOpBlock];  // new Ex, Ez of the last three planes
    __shared__ double red[8];
    const YeeArgs &a = A.a;
    const int I = a.I, J = a.J, K = a.K;
    const int tid = threadIdx.x, tx = tid % TW, ty = tid / TW;
    const int txi = tile % A.ntx, tyi = tile / A.ntx;
    const int x0 = 1 + txi * A.tw, y0 = 1 + tyi * A.th;
    const int i = x0 + tx, j = y0 + ty;
    // lanes past the halo column / row repeat it (their results are unused): loads stay inside the
    // grid and touch no line of the next tile that the halo does not
    const int ic = min(min(i, x0 + A.tw), I - 1), jc = min(min(j, y0 + A.th), J - 1);
    const bool own = tx < A.tw && ty < A.th && i <= I - 2 && j <= J - 2;
    const bool inner = i <= I - 2 && j <= J - 2;  // E follows the update (else it keeps its value)
    const int kb = 1 + kcix * A.kc, ke = min(kb + A.kc, K - 1);
    const Off P = (Off)a.P, IO = (Off)I;
    const Off p = (Off)ic + IO * (Off)jc;
    const unsigned nal4 = A.nal4;
    const __amdgpu_buffer_rsrc_t rsrc = buf_rsrc(a.Ex, 6 * nal4);
    const __amdgpu_buffer_rsrc_t rdst = buf_rsrc(A.dst, 6 * nal4);
    const __amdgpu_buffer_rsrc_t rps = buf_rsrc(A.psrc, A.pbytes);
    const __amdgpu_buffer_rsrc_t rpd = buf_rsrc(A.pdst, A.pbytes);
    const MatT *__restrict__ mat = static_cast<const MatT *>(a.mat);
    const float4 *__restrict__ coefp = reinterpret_cast<const float4 *>(A.coefp);
    const float *__restrict__ cbe = A.cbe;
    const float cbh = A.cbh;

    // CPML slabs of this thread's column and row; x / y memory offsets (floats) at plane 0
    const int sx = slab_of(ic, I, A.nlo[0], A.nhi[0]);
    const int sy = slab_of(jc, J, A.nlo[1], A.nhi[1]);
    const Off qx0 = (Off)(sx < 0 ? 0 : sx) + (Off)A.W[0] * (Off)jc, qxs = (Off)A.W[0] * (Off)J;
    const Off qy0 = (Off)ic + IO * (Off)(sy < 0 ? 0 : sy), qys = IO * (Off)A.W[1];
    const bool srcs = ke >= a.src_k0 && kb < a.src_k1;  // [kb, ke] meets a source plane

    auto body = [&](auto src_c) {
        constexpr bool SRC = decltype(src_c)::value;
        constexpr bool SX = (SL & 1) != 0, SY = (SL & 2) != 0, SZ = (SL & 4) != 0, PSI = SL != 0;
        Fac fx{}, fy{};
        if constexpr (SX) fx = fac_at(A.tab[0], I, ic);
        if constexpr (SY) fy = fac_at(A.tab[1], J, jc);
        auto zslab_at = [&](int k) { return SZ ? slab_of(k, K, A.nlo[2], A.nhi[2]) : -1; };
        // memory offsets (bytes) of this thread at plane k; out of range outside the slab
        auto pofx = [&](int k) { return sx >= 0 ? (qx0 + qxs * (Off)k) * 4u : kOob; };
        auto pofy = [&](int k) { return sy >= 0 ? (qy0 + qys * (Off)k) * 4u : kOob; };
        auto pofz = [&](int sz) { return sz >= 0 ? (p + P * (Off)sz) * 4u : kOob; };
        // plane k's fields and memories, the coefficients of material id mk (plane k's, already
        // loaded) and the id of plane kn
        auto load_fields = [&](int k, OpIn<MatT> &in) {
            const Off c4 = (p + P * (Off)k) * 4u;
            in.h[0] = buf_ld(rsrc, c4, 3u * nal4), in.h[1] = buf_ld(rsrc, c4, 4u * nal4);
            in.h[2] = buf_ld(rsrc, c4, 5u * nal4);
            in.nb[0] = buf_ld(rsrc, c4 - 4u, 4u * nal4), in.nb[1] = buf_ld(rsrc, c4 - 4u, 5u * nal4);
            in.nb[2] = buf_ld(rsrc, c4 - 4u * IO, 3u * nal4), in.nb[3] = buf_ld(rsrc, c4 - 4u * IO, 5u * nal4);
            in.eo[0] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc, (int)c4, 0, kOwnAux));
            in.eo[1] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc, (int)c4, (int)nal4, kOwnAux));
            in.eo[2] = __builtin_bit_cast(float,
                                          __builtin_amdgcn_raw_buffer_load_b32(rsrc, (int)c4, (int)(2u * nal4), kOwnAux));
            if constexpr (SX) {
                const unsigned vx = pofx(k);
                in.pe[0] = buf_ld(rps, vx, A.poE[0][0]), in.pe[1] = buf_ld(rps, vx, A.poE[0][1]);
                in.ph[0] = buf_ld(rps, vx, A.poH[0][0]), in.ph[1] = buf_ld(rps, vx, A.poH[0][1]);
            }
            if constexpr (SY) {
                const unsigned vy = pofy(k);
                in.pe[2] = buf_ld(rps, vy, A.poE[1][0]), in.pe[3] = buf_ld(rps, vy, A.poE[1][1]);
                in.ph[2] = buf_ld(rps, vy, A.poH[1][0]), in.ph[3] = buf_ld(rps, vy, A.poH[1][1]);
            }
            if constexpr (SZ) {
                const unsigned vz = pofz(zslab_at(k));
                in.pe[4] = buf_ld(rps, vz, A.poE[2][0]), in.pe[5] = buf_ld(rps, vz, A.poE[2][1]);
                in.ph[4] = buf_ld(rps, vz, A.poH[2][0]), in.ph[5] = buf_ld(rps, vz, A.poH[2][1]);
            }
        };
        auto gather = [&](OpIn<MatT> &in, int mk) {
            if constexpr (ISO) {
                in.cb = cbe[mk];
            } else {
                const float4 *cr = coefp + 3 * mk;
                in.c0 = cr[0], in.c1 = cr[1], in.c2 = cr[2];
            }
            in.m = mk;
        };
        auto load = [&](int k, int mk, int kn, OpIn<MatT> &in) {
            // the id first: the gather that needs it is issued one group later, and waiting for
            // it then must not wait for this group's field loads (the counter retires in order)
            in.mn = mat[p + P * (Off)kn];
            const Off c4 = (p + P * (Off)k) * 4u;
            in.h[0] = buf_ld(rsrc, c4, 3u * nal4), in.h[1] = buf_ld(rsrc, c4, 4u * nal4);
            in.h[2] = buf_ld(rsrc, c4, 5u * nal4);
            in.nb[0] = buf_ld(rsrc, c4 - 4u, 4u * nal4), in.nb[1] = buf_ld(rsrc, c4 - 4u, 5u * nal4);
            in.nb[2] = buf_ld(rsrc, c4 - 4u * IO, 3u * nal4), in.nb[3] = buf_ld(rsrc, c4 - 4u * IO, 5u * nal4);
            in.eo[0] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc, (int)c4, 0, kOwnAux));
            in.eo[1] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc, (int)c4, (int)nal4, kOwnAux));
            in.eo[2] = __builtin_bit_cast(float,
                                          __builtin_amdgcn_raw_buffer_load_b32(rsrc, (int)c4, (int)(2u * nal4), kOwnAux));
            if constexpr (SX) {
                const unsigned vx = pofx(k);
                in.pe[0] = buf_ld(rps, vx, A.poE[0][0]), in.pe[1] = buf_ld(rps, vx, A.poE[0][1]);
                in.ph[0] = buf_ld(rps, vx, A.poH[0][0]), in.ph[1] = buf_ld(rps, vx, A.poH[0][1]);
            }
            if constexpr (SY) {
                const unsigned vy = pofy(k);
                in.pe[2] = buf_ld(rps, vy, A.poE[1][0]), in.pe[3] = buf_ld(rps, vy, A.poE[1][1]);
                in.ph[2] = buf_ld(rps, vy, A.poH[1][0]), in.ph[3] = buf_ld(rps, vy, A.poH[1][1]);
            }
            if constexpr (SZ) {
                const unsigned vz = pofz(zslab_at(k));
                in.pe[4] = buf_ld(rps, vz, A.poE[2][0]), in.pe[5] = buf_ld(rps, vz, A.poE[2][1]);
                in.ph[4] = buf_ld(rps, vz, A.poH[2][0]), in.ph[5] = buf_ld(rps, vz, A.poH[2][1]);
            }
            if constexpr (ISO) {
                in.cb = cbe[mk];
            } else {
                const float4 *cr = coefp + 3 * mk;
                in.c0 = cr[0], in.c1 = cr[1], in.c2 = cr[2];
            }
            in.m = mk;
        };

        // new E of plane k from its inputs and the old Hx, Hy of plane k - 1; stores E and the
        // E memories of owned cells below ke
        auto e_plane = [&](int k, const OpIn<MatT> &c, float hxm, float hym, float en[3]) {
            // curl_x = dHz/dy - dHy/dz, curl_y = dHx/dz - dHz/dx, curl_z = dHy/dx - dHx/dy
            float x1 = c.h[2] - c.nb[3], x2 = c.h[1] - hym;
            float y1 = c.h[0] - hxm, y2 = c.h[2] - c.nb[1];
            float z1 = c.h[1] - c.nb[0], z2 = c.h[0] - c.nb[2];
            if constexpr (PSI) {
                // psi <- b psi + a d; d <- d / kappa + psi (0, 0, 1 outside the slabs: d unchanged)
                const bool live = own && k < ke;
                float tx1 = x1, tx2 = x2, ty1 = y1, ty2 = y2, tz1 = z1, tz2 = z2;
                if constexpr (SX) {
                    const float nx0 = fx.bE * c.pe[0] + fx.aE * y2, nx1 = fx.bE * c.pe[1] + fx.aE * z1;
                    ty2 = y2 * fx.kE + nx0, tz1 = z1 * fx.kE + nx1;
                    const unsigned vx = live ? pofx(k) : kOob;
                    buf_st(rpd, vx, A.poE[0][0], nx0), buf_st(rpd, vx, A.poE[0][1], nx1);
                }
                if constexpr (SY) {
                    const float ny0 = fy.bE * c.pe[2] + fy.aE * z2, ny1 = fy.bE * c.pe[3] + fy.aE * x1;
                    tz2 = z2 * fy.kE + ny0, tx1 = x1 * fy.kE + ny1;
                    const unsigned vy = live ? pofy(k) : kOob;
                    buf_st(rpd, vy, A.poE[1][0], ny0), buf_st(rpd, vy, A.poE[1][1], ny1);
                }
                if constexpr (SZ) {
                    const int sz = zslab_at(k);
                    const float *tz = A.tab[2];
                    const float bz = tz[k], az = tz[K + k], kz = tz[2 * K + k];  // uniform
                    const float nz0 = bz * c.pe[4] + az * x2, nz1 = bz * c.pe[5] + az * y1;
                    tx2 = x2 * kz + nz0, ty1 = y1 * kz + nz1;
                    const unsigned vz = live ? pofz(sz) : kOob;
                    buf_st(rpd, vz, A.poE[2][0], nz0), buf_st(rpd, vz, A.poE[2][1], nz1);
                }
                x1 = tx1, x2 = tx2, y1 = ty1, y2 = ty2, z1 = tz1, z2 = tz2;
            }
            // E <- caE E + cbE curl H (- cjE J); with caE = 1 the product's separate rounding
            // keeps the general form's result bit for bit
            float ex, ey, ez;
            if constexpr (ISO) {
                ex = c.eo[0] + __fmul_rn(c.cb, x1 - x2);
                ey = c.eo[1] + __fmul_rn(c.cb, y1 - y2);
                ez = c.eo[2] + __fmul_rn(c.cb, z1 - z2);
            } else {
                ex = c.c0.x * c.eo[0] + c.c0.w * (x1 - x2);
                ey = c.c0.y * c.eo[1] + c.c1.x * (y1 - y2);
                ez = c.c0.z * c.eo[2] + c.c1.y * (z1 - z2);
            }
            if constexpr (SRC) {
                float sv[3];
                if (source3(a, k, p, 0, sv)) {
                    const float *cf = a.coef + c.m * kCoefStride;
                    ex -= cf[6] * sv[0];
                    ey -= cf[7] * sv[1];
                    ez -= cf[8] * sv[2];
                }
            }
            const bool upd = inner && k <= K - 2;
            en[0] = upd ? ex : c.eo[0];
            en[1] = upd ? ey : c.eo[1];
            en[2] = upd ? ez : c.eo[2];
            const unsigned v = own && k < ke ? (p + P * (Off)k) * 4u : kOob;
            buf_st(rdst, v, 0u, en[0]);
            buf_st(rdst, v, nal4, en[1]);
            buf_st(rdst, v, 2u * nal4, en[2]);
        };

        // new H of plane k - 1 (owned cells) from the previous plane's values, this plane's new
        // Ex, Ey, and the new E of the i+1 / j+1 neighbours at plane k - 1
        auto h_plane = [&](int k1, const OpPrev &q, const float en[3], int slot) {
            const float ey_ip = lane_next(q.en[1]), ez_ip = lane_next(q.en[2]);
            const int up = ty < TH - 1 ? tid + TW : tid;  // owned rows have their j+1 row in the tile
            const float ex_jp = ring[slot][0][up], ez_jp = ring[slot][1][up];
            // curl_x = dEz/dy - dEy/dz, curl_y = dEx/dz - dEz/dx, curl_z = dEy/dx - dEx/dy
            float x1 = ez_jp - q.en[2], x2 = en[1] - q.en[1];
            float y1 = en[0] - q.en[0], y2 = ez_ip - q.en[2];
            float z1 = ey_ip - q.en[1], z2 = ex_jp - q.en[0];
            if constexpr (PSI) {
                float tx1 = x1, tx2 = x2, ty1 = y1, ty2 = y2, tz1 = z1, tz2 = z2;
                if constexpr (SX) {
                    const float nx0 = fx.bH * q.ph[0] + fx.aH * y2, nx1 = fx.bH * q.ph[1] + fx.aH * z1;
                    ty2 = y2 * fx.kH + nx0, tz1 = z1 * fx.kH + nx1;
                    const unsigned vx = own ? pofx(k1) : kOob;
                    buf_st(rpd, vx, A.poH[0][0], nx0), buf_st(rpd, vx, A.poH[0][1], nx1);
                }
                if constexpr (SY) {
                    const float ny0 = fy.bH * q.ph[2] + fy.aH * z2, ny1 = fy.bH * q.ph[3] + fy.aH * x1;
                    tz2 = z2 * fy.kH + ny0, tx1 = x1 * fy.kH + ny1;
                    const unsigned vy = own ? pofy(k1) : kOob;
                    buf_st(rpd, vy, A.poH[1][0], ny0), buf_st(rpd, vy, A.poH[1][1], ny1);
                }
                if constexpr (SZ) {
                    const int sz = zslab_at(k1);
                    const float *tz = A.tab[2];
                    const float bz = tz[3 * K + k1], az = tz[4 * K + k1], kz = tz[5 * K + k1];
                    const float nz0 = bz * q.ph[4] + az * x2, nz1 = bz * q.ph[5] + az * y1;
                    tx2 = x2 * kz + nz0, ty1 = y1 * kz + nz1;
                    const unsigned vz = own ? pofz(sz) : kOob;
                    buf_st(rpd, vz, A.poH[2][0], nz0), buf_st(rpd, vz, A.poH[2][1], nz1);
                }
                x1 = tx1, x2 = tx2, y1 = ty1, y2 = ty2, z1 = tz1, z2 = tz2;
            }
            // H <- caH H - cbH curl E (- cjH M)
            float hx, hy, hz;
            if constexpr (ISO) {
                hx = q.h[0] - __fmul_rn(cbh, x1 - x2);
                hy = q.h[1] - __fmul_rn(cbh, y1 - y2);
                hz = q.h[2] - __fmul_rn(cbh, z1 - z2);
            } else {
                hx = q.ca[0] * q.h[0] - q.cb[0] * (x1 - x2);
                hy = q.ca[1] * q.h[1] - q.cb[1] * (y1 - y2);
                hz = q.ca[2] * q.h[2] - q.cb[2] * (z1 - z2);
            }
            if constexpr (SRC) {
                float sv[3];
                if (source3(a, k1, p, 3, sv)) {
                    const float *cf = a.coef + q.m * kCoefStride;
                    hx -= cf[15] * sv[0];
                    hy -= cf[16] * sv[1];
                    hz -= cf[17] * sv[2];
                }
            }
            const unsigned v = own ? (p + P * (Off)k1) * 4u : kOob;
            buf_st(rdst, v, 3u * nal4, hx);
            buf_st(rdst, v, 4u * nal4, hy);
            buf_st(rdst, v, 5u * nal4, hz);
        };

        auto keep = [&](OpPrev &q, const OpIn<MatT> &c, const float en[3]) {
            q.en[0] = en[0], q.en[1] = en[1], q.en[2] = en[2];
            q.h[0] = c.h[0], q.h[1] = c.h[1], q.h[2] = c.h[2];
            if constexpr (!ISO) {
                q.ca[0] = c.c1.z, q.ca[1] = c.c1.w, q.ca[2] = c.c2.x;
                q.cb[0] = c.c2.y, q.cb[1] = c.c2.z, q.cb[2] = c.c2.w;
            }
            q.m = c.m;
#pragma unroll
            for (int r = 0; r < 6; r++) q.ph[r] = c.ph[r];
        };
        auto barrier = [&]() {
            // LDS-only ordering: the ring writes before the barrier, its reads after it
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            __builtin_amdgcn_s_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        };

        if constexpr (KCU > 0) {
            // every input of the chunk first (ids, then fields, then the coefficient gathers that
            // need the ids), so that each plane waits only for its own loads
            OpIn<MatT> S[KCU + 1];
            MatT mr[KCU + 1];
#pragma unroll
            for (int t = 0; t <= KCU; t++) mr[t] = mat[p + P * (Off)min(kb + t, ke)];
            const Off cb4 = (p + P * (Off)(kb - 1)) * 4u;
            const float hxm0 = buf_ld(rsrc, cb4, 3u * nal4), hym0 = buf_ld(rsrc, cb4, 4u * nal4);
#pragma unroll
            for (int t = 0; t <= KCU; t++) load_fields(min(kb + t, ke), S[t]);
#pragma unroll
            for (int t = 0; t <= KCU; t++) gather(S[t], (int)mr[t]);
            OpPrev q;
            float en[3];
            e_plane(kb, S[0], hxm0, hym0, en);
            ring[0][0][tid] = en[0];
            ring[0][1][tid] = en[2];
            keep(q, S[0], en);
            barrier();
#pragma unroll
            for (int t = 1; t <= KCU; t++) {
                const int k = kb + t;
                if (k > ke) break;  // uniform: a short last chunk
                e_plane(k, S[t], q.h[0], q.h[1], en);
                ring[t % 3][0][tid] = en[0];
                ring[t % 3][1][tid] = en[2];
                h_plane(k - 1, q, en, (t - 1) % 3);
                keep(q, S[t], en);
                barrier();
            }
            return;
        }
        // Three register sets rotate: while plane k is computed from its set, the loads of plane
        // k + 1 (issued one iteration earlier) and of plane k + 2 (issued now) are in flight.
        // Set t also carries the material id of plane t + 1, whose coefficient gather is issued
        // with plane t + 1's group.
        OpIn<MatT> SA, SB, SC;
        OpPrev q;
        float en[3];
        const Off cb4 = (p + P * (Off)(kb - 1)) * 4u;
        const float hxm0 = buf_ld(rsrc, cb4, 3u * nal4), hym0 = buf_ld(rsrc, cb4, 4u * nal4);
        load(kb, (int)mat[p + P * (Off)kb], min(kb + 1, ke), SA);
        load(min(kb + 1, ke), (int)SA.mn, min(kb + 2, ke), SB);
        // plane kb: E only
        load(min(kb + 2, ke), (int)SB.mn, min(kb + 3, ke), SC);
        e_plane(kb, SA, hxm0, hym0, en);
        int slot = 0;
        ring[slot][0][tid] = en[0];
        ring[slot][1][tid] = en[2];
        keep(q, SA, en);
        barrier();
        // planes kb+1 .. ke: E of plane k, then H of plane k - 1
        auto iter = [&](int k, const OpIn<MatT> &cur, const OpIn<MatT> &nx1, OpIn<MatT> &nx2) {
            load(min(k + 2, ke), (int)nx1.mn, min(k + 3, ke), nx2);
            e_plane(k, cur, q.h[0], q.h[1], en);
            const int ns = slot == 2 ? 0 : slot + 1;
            ring[ns][0][tid] = en[0];
            ring[ns][1][tid] = en[2];
            h_plane(k - 1, q, en, slot);
            keep(q, cur, en);
            slot = ns;
            barrier();
        };
        for (int k = kb + 1; k <= ke; k += 3) {
            iter(k, SB, SC, SA);
            if (k + 1 > ke) break;
            iter(k + 1, SC, SA, SB);
            if (k + 2 > ke) break;
            iter(k + 2, SA, SB, SC);
        }
    };
    if (srcs) body(std::true_type{});
    else body(std::false_type{});

    // overlap monitors of this step on planes of the chunk: FP64 partial of the owned cells,
    // from the new E and H this thread has just stored (timefield3d.c:3195-3224)
    for (int t = 0; t < a.nmon; t++) {
        const RunMonArg &M = a.mon[t];
        if (M.k < kb || M.k >= ke) continue;  // uniform over the block
        double re = 0.0, im = 0.0;
        if (own) {
            const float *D = A.dst;
            const Off nal = (Off)(nal4 / 4u), c = p + P * (Off)M.k;
            const double ex = ld(D, c), ey = ld(D, nal + c), hx = ld(D, 3 * nal + c), hy = ld(D, 4 * nal + c);
            double w[8];
            for (int r = 0; r < 8; r++) w[r] = ld(M.W + (long long)r * a.P, p);
            re = (hx * w[2] - hy * w[0]) + (ey * w[4] - ex * w[6]);
            im = (hx * w[3] - hy * w[1]) + (ey * w[5] - ex * w[7]);
        }
        const double2 v = sum256(re, im, red);
        double2 *part = a.step ? M.part + (long long)(*a.step - a.step0) * M.stride : M.part;
        if (tid == 0) part[tile] = v;
    }
}

// Block lb of slab class SL -> (tile, z chunk).  The tiles clear of the x slabs are
// [bx0, bx1), those clear of the y slabs [by0, by1), the chunks clear of the z slabs [bk0, bk1)
// (each a contiguous range, possibly empty).  Within the clear chunks class 0 takes the tiles clear
// of both, class 1 those in an x slab only, class 2 in a y slab only, class 3 in both; class 7
// takes every tile of the other chunks.  Enumerations run tile-fastest, so consecutive blocks (one
// XCD) are neighbouring tiles.
__device__ __forceinline__ void onepass_decode(const OnePassArgs &A, int cls, int lb, int *tile, int *kc) {
    const int nbx = A.bx1 - A.bx0, nby = A.by1 - A.by0, nbk = A.bk1 - A.bk0;
    if (cls == 7) {
        const int c = lb / A.tiles;
        *tile = lb % A.tiles;
        *kc = c < A.bk0 ? c : c + nbk;
        return;
    }
    const int cx = (cls & 1) ? A.ntx - nbx : nbx, cy = (cls & 2) ? A.nty - nby : nby;
    const int per = cx * cy, t = lb % per, ux = t % cx, uy = t / cx;
    *kc = A.bk0 + lb / per;
    const int tx = (cls & 1) ? (ux < A.bx0 ? ux : ux + nbx) : A.bx0 + ux;
    const int ty = (cls & 2) ? (uy < A.by0 ? uy : uy + nby) : A.by0 + uy;
    *tile = tx + A.ntx * ty;
}

template <typename MatT, int TW, bool NT, bool ISO, int SL, int KCU>
__global__ __launch_bounds__(kOpBlock) void k_onepass(OnePassArgs A) {
    const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
    int tile, kc;
    onepass_decode(A, SL, lb, &tile, &kc);
    onepass_block<MatT, TW, NT, ISO, SL, KCU>(A, tile, kc);
}

}  // namespace

bool onepass_supported(const fdtd_dev *d) {
    const Cpml &cp = d->cpml;
    const long long pmem = 4LL * 4 * (cp.psi_n[0] + cp.psi_n[1] + cp.psi_n[2]);
    return cp.on && cp.psi_base && !d->dense_src && d->mat_bytes <= 2 && d->I >= 3 && d->J >= 3 && d->K >= 3 &&
           24LL * d->Nal < (1LL << 31) && pmem < (1LL << 31) && (long long)d->src.size() <= kMaxSrcPlanes;
}

int onepass_plane_slots(const fdtd_dev *d, int tw) {
    const int th = kOpBlock / tw;
    return ((d->I - 2 + tw - 2) / (tw - 1)) * ((d->J - 2 + th - 2) / (th - 1));  // as onepass_fill
}

int onepass_default_chunk(const fdtd_dev *d) {
    (void)d;
    return 4;  // the chunk whose loads are all issued up front (fastest measured: 4 < 8 < 16 planes)
}

int onepass_default_width(const fdtd_dev *d) {
    (void)d;
    return 64;
}

int onepass_fill(fdtd_dev *d, double dt, int tw, int kc, OnePassArgs *Ap) {
    if (tw != 64) {
        set_error("single-pass step: tile width must be 64");
        return -2;
    }
    if (kc < 1) {
        set_error("single-pass step: z planes per chunk must be positive");
        return -2;
    }
    OnePassArgs &A = *Ap;
    A = OnePassArgs{};
    PmlArgs pm;
    int rc = fill_step_args(d, dt, &A.a, &pm);
    if (rc) return rc;
    YeeArgs &a = A.a;
    a.nrs = d->n_run_src;
    for (int q = 0; q < d->n_run_src; q++) a.rs[q] = d->run_src[q];
    set_source_range(a);
    a.nmon = d->n_run_mon;
    for (int q = 0; q < a.nmon; q++) a.mon[q] = d->run_mon[q];
    const Cpml &cp = d->cpml;
    const long long pn = cp.psi_n[0] + cp.psi_n[1] + cp.psi_n[2];
    for (int ax = 0; ax < 3; ax++) {
        A.tab[ax] = cp.tab + cp.tab_off[ax];
        A.nlo[ax] = cp.nlo[ax];
        A.nhi[ax] = cp.nhi[ax];
        A.W[ax] = cp.W[ax];
        for (int q = 0; q < 2; q++) {
            A.poE[ax][q] = cp.psi[0][ax][q] ? (unsigned)(4 * (cp.psi[0][ax][q] - cp.psi_base)) : 0u;
            A.poH[ax][q] = cp.psi[1][ax][q] ? (unsigned)(4 * (cp.psi[1][ax][q] - cp.psi_base)) : 0u;
        }
    }
    A.coefp = d->d_coef + coef_packed_offset(d->nmat);
    A.cbe = d->d_coef + coef_iso_offset(d->nmat);
    A.cbh = d->coef_cbh;
    A.iso = d->coef_iso ? 1 : 0;
    A.psrc = cp.psi_base;
    A.pdst = cp.psi_alt;
    A.pbytes = (unsigned)(16 * pn);
    A.dst = d->alt_base;
    A.nal4 = (unsigned)(4 * d->Nal);
    const int th = kOpBlock / tw;
    A.ntx = (d->I - 2 + tw - 2) / (tw - 1);
    A.nty = (d->J - 2 + th - 2) / (th - 1);
    A.tw = (d->I - 2 + A.ntx - 1) / A.ntx;  // owned columns / rows per tile, at most tw - 1 / th - 1
    A.th = (d->J - 2 + A.nty - 1) / A.nty;
    A.tiles = A.ntx * A.nty;
    A.kc = kc;
    A.nkc = (d->K - 2 + kc - 1) / kc;
    A.vb0 = 0;
    // the box of tiles and chunks clear of every slab (the kernel's own predicates, evaluated per
    // tile / chunk; slabs sit at the ends of each axis, so the clear ones are a contiguous range)
    auto clear_range = [](int n, auto touches, int *lo, int *hi) {
        *lo = 0;
        while (*lo < n && touches(*lo)) ++*lo;
        *hi = *lo;
        while (*hi < n && !touches(*hi)) ++*hi;
        if (*hi == *lo) *lo = *hi = 0;
    };
    const int I = d->I, J = d->J, K = d->K;
    clear_range(A.ntx, [&](int t) {
        const int x0 = 1 + t * A.tw, xh = std::min(x0 + A.tw, I - 1);
        return x0 < A.nlo[0] || (A.nhi[0] > 0 && xh >= I - 1 - A.nhi[0]);
    }, &A.bx0, &A.bx1);
    clear_range(A.nty, [&](int t) {
        const int y0 = 1 + t * A.th, yh = std::min(y0 + A.th, J - 1);
        return y0 < A.nlo[1] || (A.nhi[1] > 0 && yh >= J - 1 - A.nhi[1]);
    }, &A.by0, &A.by1);
    clear_range(A.nkc, [&](int c) {
        const int kb = 1 + c * kc, ke = std::min(kb + kc, K - 1);
        return kb - 1 < A.nlo[2] || (A.nhi[2] > 0 && ke >= K - 1 - A.nhi[2]);
    }, &A.bk0, &A.bk1);
    const int nbx = A.bx1 - A.bx0, nby = A.by1 - A.by0, nbk = A.bk1 - A.bk0;
    A.ncls[0] = nbx * nby * nbk;
    A.ncls[1] = (A.ntx - nbx) * nby * nbk;
    A.ncls[2] = nbx * (A.nty - nby) * nbk;
    A.ncls[3] = (A.ntx - nbx) * (A.nty - nby) * nbk;
    A.ncls[4] = (A.nkc - nbk) * A.tiles;
    return 0;
}

namespace {

// The launches of one step, one per non-empty slab class (class 7 = all memories first); t0 / t1
// bracket them all.
template <typename MatT, int TW, bool NT, bool ISO, int KCU>
void launch_classes(hipStream_t st, hipEvent_t t0, hipEvent_t t1, const OnePassArgs &A) {
    static constexpr int order[5] = {4, 3, 1, 2, 0};  // ncls index: z chunks, xy, x, y, clear
    int first = -1, last = -1;
    for (int q = 0; q < 5; q++)
        if (A.ncls[order[q]] > 0) {
            if (first < 0) first = q;
            last = q;
        }
    for (int q = 0; q < 5; q++) {
        const int c = order[q];
        if (A.ncls[c] <= 0) continue;
        const dim3 grid((unsigned)A.ncls[c]), block(kOpBlock);
        hipEvent_t a0 = q == first ? t0 : nullptr, a1 = q == last ? t1 : nullptr;
        switch (c) {
            case 0: hipExtLaunchKernelGGL((k_onepass<MatT, TW, NT, ISO, 0, KCU>), grid, block, 0, st, a0, a1, 0, A); break;
            case 1: hipExtLaunchKernelGGL((k_onepass<MatT, TW, NT, ISO, 1, KCU>), grid, block, 0, st, a0, a1, 0, A); break;
            case 2: hipExtLaunchKernelGGL((k_onepass<MatT, TW, NT, ISO, 2, KCU>), grid, block, 0, st, a0, a1, 0, A); break;
            case 3: hipExtLaunchKernelGGL((k_onepass<MatT, TW, NT, ISO, 3, KCU>), grid, block, 0, st, a0, a1, 0, A); break;
            default: hipExtLaunchKernelGGL((k_onepass<MatT, TW, NT, ISO, 7, KCU>), grid, block, 0, st, a0, a1, 0, A); break;
        }
    }
}

template <typename MatT, bool NT, bool ISO>
void launch_tw(int tw, hipStream_t st, hipEvent_t t0, hipEvent_t t1, const OnePassArgs &A) {
    (void)tw;
    if (A.kc == 4) launch_classes<MatT, 64, NT, ISO, 4>(st, t0, t1, A);
    else launch_classes<MatT, 64, NT, ISO, 0>(st, t0, t1, A);
}

template <typename MatT, bool NT>
void launch_iso(int tw, hipStream_t st, hipEvent_t t0, hipEvent_t t1, const OnePassArgs &A) {
    if (A.iso) launch_tw<MatT, NT, true>(tw, st, t0, t1, A);
    else launch_tw<MatT, NT, false>(tw, st, t0, t1, A);
}

}  // namespace

int onepass_step(fdtd_dev *d, double dt, int tw, int kc) {
    fields_changed(d);
    int rc = fused_prepare(d);
    if (rc) return rc;
    OnePassArgs A;
    if ((rc = onepass_fill(d, dt, tw, kc, &A))) return rc;
    if (d->n_run_mon > 0) d->n_run_mon = 0;  // consumed
    const bool u8 = d->mat_bytes == 1;
    const bool nt = streaming_loads(d->N, d->stream_mode);
    hipEvent_t t0, t1;
    timing_events(d, 2, &t0, &t1);
    if (u8 && nt) launch_iso<uint8_t, true>(tw, d->stream, t0, t1, A);
    else if (u8) launch_iso<uint8_t, false>(tw, d->stream, t0, t1, A);
    else if (nt) launch_iso<uint16_t, true>(tw, d->stream, t0, t1, A);
    else launch_iso<uint16_t, false>(tw, d->stream, t0, t1, A);
    FDTD_TRY(hipGetLastError());
    fused_swap(d);
    return 0;
}

}  // namespace fdtd
