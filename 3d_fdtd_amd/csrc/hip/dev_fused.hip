// Fused E+H step for the PML (PEC-walled) boundary: one kernel per time step, each field
// component read once and written once.  Reference semantics of the update:
// timefield3d.c:1679-1837 (semi-implicit E/H updates); the PML terms are the extension described
// in oracle/fdtd_oracle.c.
//
// Mapping (gfx950, 64-wide wavefronts):
//   * a wavefront owns a strip of 62 x-columns, plus one halo lane on each side (lane 0 = column
//     x0-1, lane 63 = column x0+62), so the i-1 / i+1 neighbours are lane shuffles and no data
//     crosses wavefronts;
//   * each lane owns TY consecutive y rows plus a halo row below (H only) and above (E only),
//     kept in registers;
//   * the wavefront marches a z chunk upward: at plane k it computes E_new(k) from H_old(k),
//     H_old(k-1) (registers), then H_new(k-1) from E_new(k-1) (registers) and E_new(k); the
//     inputs of plane k+1 are loaded while plane k is computed.
// Halo values are recomputed from the same old inputs with the same arithmetic, so they equal
// the owner's values.  Fields and PML memories are ping-ponged (read src, write dst), so the
// redundant halo work never races with an owner's store.  Boundary cells (index 0 and N-1 on
// any axis) are never written: both buffers hold the same wall values.
//
// Arguments are kept compact (two field base pointers and one component stride, one PML memory
// base with per-axis offsets, 32-bit cell indices) so that the uniform state fits the scalar
// register file; the launcher falls back to the split kernels when a grid needs more (more than
// two source planes, a dense source volume, or more than 2^31 addressable floats).
#include <cmath>

#include "dev_internal.h"

namespace fdtd {
namespace {

constexpr int kStrip = 62;  // owned columns per wavefront

struct FusedArgs {
    const float *src;  // 6 component arrays, stride nal
    float *dst;
    int nal;
    const void *mat;
    const float *coef;
    int I, J, K, P;
    int nwx, nty, nkc, kc_len;
    int sk[2];              // source planes (-1 = none); values at J[m * P + p]
    const float *sJ[2];
    int nrs;                // tabulated run sources (override the plane inside their window)
    RunSrcArg rs[kMaxRunSrc];
    // PML: tables per axis at tab + toff[a] (bE,aE,kE,bH,aH,kH x N_a); memory psi(a, f, q) at
    // pbase + poff[a] + (2 f + q) * pn[a]
    const float *tab;
    int toff[3];
    const float *psrc;
    float *pdst;
    int poff[3], pn[3];
    int nlo[3], nhi[3], W[3];
};

template <typename MatT>
__device__ __forceinline__ int mat_id(const void *mat, int c) {
    return (int)static_cast<const MatT *>(mat)[c];
}

// Source components m0..m0+2 of plane cell p on plane k (Jp = stored plane or null).
__device__ __forceinline__ bool fsource3(const FusedArgs &a, const float *Jp, int k, int p, int m0, float v[3]) {
    for (int q = 0; q < a.nrs; q++) {
        const RunSrcArg &r = a.rs[q];
        if (r.k == k && r.win[p]) {
            for (int m = 0; m < 3; m++) v[m] = r.c1 * r.PR[(m0 + m) * a.P + p] + r.c2 * r.PI[(m0 + m) * a.P + p];
            return true;
        }
    }
    if (!Jp) return false;
    for (int m = 0; m < 3; m++) v[m] = Jp[(m0 + m) * a.P + p];
    return true;
}

__device__ __forceinline__ int slab_index(int n, int N, int nlo, int nhi) {
    if (n < nlo) return n;
    if (nhi > 0 && n >= N - 1 - nhi) return nlo + n - (N - 1 - nhi);
    return -1;
}

// psi_new = b psi_old + a d; returns d/kappa + psi_new; stores psi_new when `own`
__device__ __forceinline__ float pml_rw(const FusedArgs &A, int q, float b, float a, float kinv, float d, bool own) {
    float v = b * A.psrc[q] + a * d;
    if (own) A.pdst[q] = v;
    return d * kinv + v;
}

__device__ __forceinline__ float shfl_up1(float v) { return __shfl_up(v, 1, 64); }
__device__ __forceinline__ float shfl_dn1(float v) { return __shfl_down(v, 1, 64); }

template <typename MatT, int TY, bool PML>
__global__ __launch_bounds__(256) void k_step_fused(FusedArgs a) {
    const int lane = threadIdx.x & 63;
    const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const int wid = lb * 4 + (threadIdx.x >> 6);
    if (wid >= a.nwx * a.nty * a.nkc) return;
    const int wx = wid % a.nwx;
    const int rest = wid / a.nwx;
    const int ty = rest % a.nty;
    const int kc = rest / a.nty;

    const int i = kStrip * wx + lane;  // lane 0 is column 62*wx (halo), lanes 1..62 owned
    const bool col_int = i >= 1 && i <= a.I - 2;
    const bool col_own = lane >= 1 && lane <= kStrip && col_int;
    const int ic = i <= a.I - 1 ? i : a.I - 1;  // clamped column for address formation
    const int j0 = 1 + TY * ty;
    const int k0 = 1 + a.kc_len * kc;
    const int k1 = min(k0 + a.kc_len, a.K - 1);  // planes [k0, k1) are owned
    const int nal = a.nal;
    const float *__restrict__ S = a.src;
    float *__restrict__ D = a.dst;

    int jr[TY + 2];
    bool row_int[TY + 2];
#pragma unroll
    for (int r = 0; r < TY + 2; r++) {
        int j = j0 - 1 + r;
        row_int[r] = j >= 1 && j <= a.J - 2;
        jr[r] = j <= a.J - 1 ? j : a.J - 1;
    }

    // PML per-column values
    int sx = -1;
    float bxE = 0, axE = 0, kxE = 1, bxH = 0, axH = 0, kxH = 1;
    if (PML) {
        const float *tx = a.tab + a.toff[0];
        sx = slab_index(ic, a.I, a.nlo[0], a.nhi[0]);
        bxE = tx[ic];
        axE = tx[a.I + ic];
        kxE = tx[2 * a.I + ic];
        bxH = tx[3 * a.I + ic];
        axH = tx[4 * a.I + ic];
        kxH = tx[5 * a.I + ic];
    }

    float hx[TY + 2], hy[TY + 2], hz[TY + 2];
    float pxh[TY + 2], pyh[TY + 2], pzh[TY + 2];
    float enx[TY + 2], eny[TY + 2], enz[TY + 2];
    float epx[TY + 2], epy[TY + 2], epz[TY + 2];
    int mprev[TY + 2];
    {
        const int base = (k0 - 1) * a.P + ic;
#pragma unroll
        for (int r = 1; r < TY + 2; r++) {
            const int c = base + jr[r] * a.I;
            pxh[r] = S[3 * nal + c];
            pyh[r] = S[4 * nal + c];
            pzh[r] = 0.0f;
            epx[r] = epy[r] = epz[r] = 0.0f;
            mprev[r] = 0;
        }
    }

    // inputs of the current plane (loaded at the top of each iteration: the many resident
    // wavefronts hide the latency better than a register-hungry prefetch of plane k+1)
    float nhx[TY + 2], nhy[TY + 2], nhz[TY + 2], nex[TY + 2], ney[TY + 2], nez[TY + 2];
    int nmt[TY + 2];
    auto load_plane = [&](int kk) {
        const int pb = kk * a.P + ic;
#pragma unroll
        for (int r = 0; r < TY + 2; r++) {
            const int c = pb + jr[r] * a.I;
            nhx[r] = S[3 * nal + c];
            nhy[r] = S[4 * nal + c];
            nhz[r] = S[5 * nal + c];
            if (r > 0) {
                nex[r] = S[c];
                ney[r] = S[nal + c];
                nez[r] = S[2 * nal + c];
                nmt[r] = mat_id<MatT>(a.mat, c);
            }
        }
    };
    float eox[TY + 2], eoy[TY + 2], eoz[TY + 2];
    for (int k = k0; k <= k1; k++) {
        load_plane(k);
        const int base = k * a.P + ic;
        const bool plane_int = k <= a.K - 2;
        const bool plane_own = k < k1;
        int mcur[TY + 2];
#pragma unroll
        for (int r = 0; r < TY + 2; r++) {
            hx[r] = nhx[r];
            hy[r] = nhy[r];
            hz[r] = nhz[r];
            if (r > 0) {
                eox[r] = nex[r];
                eoy[r] = ney[r];
                eoz[r] = nez[r];
                mcur[r] = nmt[r];
            }
        }
        const float *Jp = k == a.sk[0] ? a.sJ[0] : (k == a.sk[1] ? a.sJ[1] : nullptr);
        int sz = -1;
        float bzE = 0, azE = 0, kzE = 1;
        if (PML) {
            const float *tz = a.tab + a.toff[2];
            sz = slab_index(k, a.K, a.nlo[2], a.nhi[2]);
            bzE = tz[k];
            azE = tz[a.K + k];
            kzE = tz[2 * a.K + k];
        }
        // ---- E_new(k), rows 1..TY+1 ----
#pragma unroll
        for (int r = 1; r < TY + 2; r++) {
            const int c = base + jr[r] * a.I;
            const float hy_im = shfl_up1(hy[r]), hz_im = shfl_up1(hz[r]);
            float x1 = hz[r] - hz[r - 1], x2 = hy[r] - pyh[r];
            float y1 = hx[r] - pxh[r], y2 = hz[r] - hz_im;
            float z1 = hy[r] - hy_im, z2 = hx[r] - hx[r - 1];
            const bool upd = col_int && row_int[r] && plane_int;
            const bool own = col_own && r <= TY && row_int[r] && plane_own && plane_int;
            if (PML) {
                const int j = jr[r];
                const float *ty_ = a.tab + a.toff[1];
                const int sy = slab_index(j, a.J, a.nlo[1], a.nhi[1]);
                const float byE = ty_[j], ayE = ty_[a.J + j], kyE = ty_[2 * a.J + j];
                float tx1 = x1 * kyE, tx2 = x2 * kzE, ty1 = y1 * kzE, ty2 = y2 * kxE, tz1 = z1 * kxE, tz2 = z2 * kyE;
                if (upd && sx >= 0) {
                    const int q = a.poff[0] + sx + a.W[0] * (j + a.J * k);
                    ty2 = pml_rw(a, q, bxE, axE, kxE, y2, own);
                    tz1 = pml_rw(a, q + a.pn[0], bxE, axE, kxE, z1, own);
                }
                if (upd && sy >= 0) {
                    const int q = a.poff[1] + ic + a.I * (sy + a.W[1] * k);
                    tz2 = pml_rw(a, q, byE, ayE, kyE, z2, own);
                    tx1 = pml_rw(a, q + a.pn[1], byE, ayE, kyE, x1, own);
                }
                if (upd && sz >= 0) {
                    const int q = a.poff[2] + ic + a.I * j + a.P * sz;
                    tx2 = pml_rw(a, q, bzE, azE, kzE, x2, own);
                    ty1 = pml_rw(a, q + a.pn[2], bzE, azE, kzE, y1, own);
                }
                x1 = tx1, x2 = tx2, y1 = ty1, y2 = ty2, z1 = tz1, z2 = tz2;
            }
            const float ex_o = eox[r], ey_o = eoy[r], ez_o = eoz[r];
            float ex = ex_o, ey = ey_o, ez = ez_o;
            if (upd) {
                const float *cf = a.coef + mcur[r] * kCoefStride;
                ex = cf[0] * ex_o + cf[3] * (x1 - x2);
                ey = cf[1] * ey_o + cf[4] * (y1 - y2);
                ez = cf[2] * ez_o + cf[5] * (z1 - z2);
                float sv[3];
                if (fsource3(a, Jp, k, c - k * a.P, 0, sv)) {
                    ex -= cf[6] * sv[0];
                    ey -= cf[7] * sv[1];
                    ez -= cf[8] * sv[2];
                }
            }
            enx[r] = ex;
            eny[r] = ey;
            enz[r] = ez;
            if (own) {
                D[c] = ex;
                D[nal + c] = ey;
                D[2 * nal + c] = ez;
            }
        }
        // ---- H_new(k-1), rows 1..TY ----
        if (k > k0) {
            const int kh = k - 1;
            const int hbase = kh * a.P + ic;
            const float *Mp = kh == a.sk[0] ? a.sJ[0] : (kh == a.sk[1] ? a.sJ[1] : nullptr);
            int szh = -1;
            float bzH = 0, azH = 0, kzH = 1;
            if (PML) {
                const float *tz = a.tab + a.toff[2];
                szh = slab_index(kh, a.K, a.nlo[2], a.nhi[2]);
                bzH = tz[3 * a.K + kh];
                azH = tz[4 * a.K + kh];
                kzH = tz[5 * a.K + kh];
            }
#pragma unroll
            for (int r = 1; r < TY + 1; r++) {
                const float ey_ip = shfl_dn1(epy[r]), ez_ip = shfl_dn1(epz[r]);
                float x1 = epz[r + 1] - epz[r], x2 = eny[r] - epy[r];
                float y1 = enx[r] - epx[r], y2 = ez_ip - epz[r];
                float z1 = ey_ip - epy[r], z2 = epx[r + 1] - epx[r];
                const bool upd = col_int && row_int[r];
                const bool own = col_own && row_int[r];
                const int c = hbase + jr[r] * a.I;
                if (PML) {
                    const int j = jr[r];
                    const float *ty_ = a.tab + a.toff[1];
                    const int sy = slab_index(j, a.J, a.nlo[1], a.nhi[1]);
                    const float byH = ty_[3 * a.J + j], ayH = ty_[4 * a.J + j], kyH = ty_[5 * a.J + j];
                    float tx1 = x1 * kyH, tx2 = x2 * kzH, ty1 = y1 * kzH, ty2 = y2 * kxH, tz1 = z1 * kxH,
                          tz2 = z2 * kyH;
                    if (upd && sx >= 0) {
                        const int q = a.poff[0] + 2 * a.pn[0] + sx + a.W[0] * (j + a.J * kh);
                        ty2 = pml_rw(a, q, bxH, axH, kxH, y2, own);
                        tz1 = pml_rw(a, q + a.pn[0], bxH, axH, kxH, z1, own);
                    }
                    if (upd && sy >= 0) {
                        const int q = a.poff[1] + 2 * a.pn[1] + ic + a.I * (sy + a.W[1] * kh);
                        tz2 = pml_rw(a, q, byH, ayH, kyH, z2, own);
                        tx1 = pml_rw(a, q + a.pn[1], byH, ayH, kyH, x1, own);
                    }
                    if (upd && szh >= 0) {
                        const int q = a.poff[2] + 2 * a.pn[2] + ic + a.I * j + a.P * szh;
                        tx2 = pml_rw(a, q, bzH, azH, kzH, x2, own);
                        ty1 = pml_rw(a, q + a.pn[2], bzH, azH, kzH, y1, own);
                    }
                    x1 = tx1, x2 = tx2, y1 = ty1, y2 = ty2, z1 = tz1, z2 = tz2;
                }
                if (own) {
                    const float *cf = a.coef + mprev[r] * kCoefStride;
                    float hxn = cf[9] * pxh[r] - cf[12] * (x1 - x2);
                    float hyn = cf[10] * pyh[r] - cf[13] * (y1 - y2);
                    float hzn = cf[11] * pzh[r] - cf[14] * (z1 - z2);
                    float sv[3];
                    if (fsource3(a, Mp, kh, c - kh * a.P, 3, sv)) {
                        hxn -= cf[15] * sv[0];
                        hyn -= cf[16] * sv[1];
                        hzn -= cf[17] * sv[2];
                    }
                    D[3 * nal + c] = hxn;
                    D[4 * nal + c] = hyn;
                    D[5 * nal + c] = hzn;
                }
            }
        }
        // carry plane k -> k-1
#pragma unroll
        for (int r = 1; r < TY + 2; r++) {
            pxh[r] = hx[r];
            pyh[r] = hy[r];
            pzh[r] = hz[r];
            epx[r] = enx[r];
            epy[r] = eny[r];
            epz[r] = enz[r];
            mprev[r] = mcur[r];
        }
    }
}

// ---------------------------------------------------------------------------------------------
// LDS-tiled fused step.  A 256-thread block owns a tile of kTileOwn x-columns by TJ y-rows and
// marches a z chunk.  Per plane k: the block gathers H_old(k) over the tile plus its i-1 / j-1
// halo into LDS (one coalesced pass, every value read once), computes E_new(k) over the tile plus
// one halo column and row (the halo is what the H update needs from the neighbours), keeps
// E_new(k) in LDS and finishes H_new(k-1) from E_new(k-1), E_new(k) and H_old(k-1).  Lanes are
// columns; wave w handles rows w, w+4, ... .  Halo E values are recomputed from the same old
// inputs with the same arithmetic as their owner's, and fields / PML memories are ping-ponged, so
// halo work never races with an owner's store.
constexpr int kTileW = 64;    // E columns of a tile (lanes)
constexpr int kTileOwn = 63;  // owned columns (lane 63 is the i+1 halo)

template <int TJ>
struct TileGeom {
    static constexpr int HR = TJ + 2;           // H rows j0-1 .. j0+TJ
    static constexpr int HC = kTileW + 1;       // H columns i0-1 .. i0+63
    static constexpr int NH = 3 * HR * HC;      // gathered H floats per plane
    static constexpr int NHL = (NH + 255) / 256;
    static constexpr int ER = TJ + 1;           // E rows j0 .. j0+TJ
    static constexpr int NE = (ER + 3) / 4;     // E rows per thread
};

template <typename MatT, int TJ, bool PML>
__global__ __launch_bounds__(256) void k_step_tiled(FusedArgs a) {
    using G = TileGeom<TJ>;
    // H_old of planes k-2, k-1, k (ring of three, [comp][row][col]); E_new of planes k-1, k
    __shared__ float Hs[3][G::NH];
    __shared__ float Es[2][3][G::ER][kTileW];
    const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const int tx = lb % a.nwx, rest = lb / a.nwx;
    const int ty = rest % a.nty, kc = rest / a.nty;
    if (kc >= a.nkc) return;  // uniform over the block
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i0 = 1 + kTileOwn * tx, j0 = 1 + TJ * ty;
    const int kb = 1 + a.kc_len * kc, ke = min(kb + a.kc_len, a.K - 1);  // owned planes [kb, ke)
    const int i = i0 + lane;
    const int ic = i <= a.I - 1 ? i : a.I - 1;
    const bool col_int = i <= a.I - 2;
    const bool col_own = lane < kTileOwn && col_int;
    const int lane_ip = lane < kTileW - 1 ? lane + 1 : lane;  // i+1 neighbour (lane 63 never owns H)
    const int nal = a.nal;
    const float *__restrict__ S = a.src;
    float *__restrict__ D = a.dst;

    // plane-relative source offsets of this thread's share of the H gather
    int hoff[G::NHL];
#pragma unroll
    for (int t = 0; t < G::NHL; t++) {
        const int q0 = threadIdx.x + 256 * t;
        const int q = q0 < G::NH ? q0 : G::NH - 1;
        const int comp = q / (G::HR * G::HC), rem = q - comp * (G::HR * G::HC);
        const int rr = rem / G::HC, cc = rem - rr * G::HC;
        const int gi = min(i0 - 1 + cc, a.I - 1), gj = min(j0 - 1 + rr, a.J - 1);
        hoff[t] = (3 + comp) * nal + gi + a.I * gj;
    }
    auto gather = [&](int kk, float *dst) {
        const int pk = kk * a.P;
        float hv[G::NHL];
#pragma unroll
        for (int t = 0; t < G::NHL; t++) hv[t] = S[hoff[t] + pk];
#pragma unroll
        for (int t = 0; t < G::NHL; t++)
            if (threadIdx.x + 256 * t < G::NH) dst[threadIdx.x + 256 * t] = hv[t];
    };
    int sx = -1;
    float bxE = 0, axE = 0, kxE = 1, bxH = 0, axH = 0, kxH = 1;
    if (PML) {
        const float *t = a.tab + a.toff[0];
        sx = slab_index(ic, a.I, a.nlo[0], a.nhi[0]);
        bxE = t[ic], axE = t[a.I + ic], kxE = t[2 * a.I + ic];
        bxH = t[3 * a.I + ic], axH = t[4 * a.I + ic], kxH = t[5 * a.I + ic];
    }
    int bm = (kb - 1) % 3;  // ring slot of plane k-1
    gather(kb - 1, Hs[bm]);
    for (int k = kb; k <= ke; k++) {
        const int bk = bm == 2 ? 0 : bm + 1;  // ring slot of plane k
        const int pk = k * a.P;
        gather(k, Hs[bk]);
        __syncthreads();
        const float *Hk = Hs[bk];
        const float *Hm = Hs[bm];
        // ---- E_new(k) over the tile + halo ----
        const bool plane_int = k <= a.K - 2, plane_own = k < ke;
        const float *Jp = k == a.sk[0] ? a.sJ[0] : (k == a.sk[1] ? a.sJ[1] : nullptr);
        int sz = -1;
        float bzE = 0, azE = 0, kzE = 1;
        if (PML) {
            const float *t = a.tab + a.toff[2];
            sz = slab_index(k, a.K, a.nlo[2], a.nhi[2]);
            bzE = t[k], azE = t[a.K + k], kzE = t[2 * a.K + k];
        }
        float (*Ek)[G::ER][kTileW] = Es[k & 1];
#pragma unroll 1
        for (int r = w; r < G::ER; r += 4) {
            const int jj = j0 + r, j = jj <= a.J - 1 ? jj : a.J - 1;
            const bool row_int = jj <= a.J - 2;
            const int c = pk + ic + a.I * j;
            const float eox = S[c], eoy = S[nal + c], eoz = S[2 * nal + c];
            const int m = mat_id<MatT>(a.mat, c);
            const int h0 = (r + 1) * G::HC + lane + 1;  // centre of row r in the gather layout
            const float hx = Hk[h0], hy = Hk[G::HR * G::HC + h0], hz = Hk[2 * G::HR * G::HC + h0];
            const float hy_im = Hk[G::HR * G::HC + h0 - 1], hz_im = Hk[2 * G::HR * G::HC + h0 - 1];
            const float hx_jm = Hk[h0 - G::HC], hz_jm = Hk[2 * G::HR * G::HC + h0 - G::HC];
            const float hx_km = Hm[h0], hy_km = Hm[G::HR * G::HC + h0];
            float x1 = hz - hz_jm, x2 = hy - hy_km;
            float y1 = hx - hx_km, y2 = hz - hz_im;
            float z1 = hy - hy_im, z2 = hx - hx_jm;
            const bool upd = col_int && row_int && plane_int;
            const bool own = col_own && r < TJ && row_int && plane_own && plane_int;
            if (PML) {
                const float *t = a.tab + a.toff[1];
                const int sy = slab_index(j, a.J, a.nlo[1], a.nhi[1]);
                const float byE = t[j], ayE = t[a.J + j], kyE = t[2 * a.J + j];
                float tx1 = x1 * kyE, tx2 = x2 * kzE, ty1 = y1 * kzE, ty2 = y2 * kxE, tz1 = z1 * kxE, tz2 = z2 * kyE;
                if (upd && sx >= 0) {
                    const int q = a.poff[0] + sx + a.W[0] * (j + a.J * k);
                    ty2 = pml_rw(a, q, bxE, axE, kxE, y2, own);
                    tz1 = pml_rw(a, q + a.pn[0], bxE, axE, kxE, z1, own);
                }
                if (upd && sy >= 0) {
                    const int q = a.poff[1] + ic + a.I * (sy + a.W[1] * k);
                    tz2 = pml_rw(a, q, byE, ayE, kyE, z2, own);
                    tx1 = pml_rw(a, q + a.pn[1], byE, ayE, kyE, x1, own);
                }
                if (upd && sz >= 0) {
                    const int q = a.poff[2] + ic + a.I * j + a.P * sz;
                    tx2 = pml_rw(a, q, bzE, azE, kzE, x2, own);
                    ty1 = pml_rw(a, q + a.pn[2], bzE, azE, kzE, y1, own);
                }
                x1 = tx1, x2 = tx2, y1 = ty1, y2 = ty2, z1 = tz1, z2 = tz2;
            }
            float ex = eox, ey = eoy, ez = eoz;
            if (upd) {
                const float *cf = a.coef + m * kCoefStride;
                ex = cf[0] * eox + cf[3] * (x1 - x2);
                ey = cf[1] * eoy + cf[4] * (y1 - y2);
                ez = cf[2] * eoz + cf[5] * (z1 - z2);
                float sv[3];
                if (fsource3(a, Jp, k, c - pk, 0, sv)) {
                    ex -= cf[6] * sv[0];
                    ey -= cf[7] * sv[1];
                    ez -= cf[8] * sv[2];
                }
            }
            Ek[0][r][lane] = ex;
            Ek[1][r][lane] = ey;
            Ek[2][r][lane] = ez;
            if (own) {
                D[c] = ex;
                D[nal + c] = ey;
                D[2 * nal + c] = ez;
            }
        }
        __syncthreads();
        // ---- H_new(k-1) over the owned cells ----
        if (k > kb) {
            const int kh = k - 1, ph = kh * a.P;
            const float *Mp = kh == a.sk[0] ? a.sJ[0] : (kh == a.sk[1] ? a.sJ[1] : nullptr);
            float (*Ep)[G::ER][kTileW] = Es[kh & 1];
            int szh = -1;
            float bzH = 0, azH = 0, kzH = 1;
            if (PML) {
                const float *t = a.tab + a.toff[2];
                szh = slab_index(kh, a.K, a.nlo[2], a.nhi[2]);
                bzH = t[3 * a.K + kh], azH = t[4 * a.K + kh], kzH = t[5 * a.K + kh];
            }
#pragma unroll 1
            for (int r = w; r < TJ; r += 4) {
                const int jj = j0 + r;
                if (!col_own || jj > a.J - 2) continue;
                const int j = jj;
                const int c = ph + ic + a.I * j;
                const int m = mat_id<MatT>(a.mat, c);
                const int h0 = (r + 1) * G::HC + lane + 1;
                const float hxo = Hm[h0], hyo = Hm[G::HR * G::HC + h0], hzo = Hm[2 * G::HR * G::HC + h0];
                const float ex = Ep[0][r][lane], ey = Ep[1][r][lane], ez = Ep[2][r][lane];
                const float ey_ip = Ep[1][r][lane_ip], ez_ip = Ep[2][r][lane_ip];
                const float ex_jp = Ep[0][r + 1][lane], ez_jp = Ep[2][r + 1][lane];
                const float ex_kp = Ek[0][r][lane], ey_kp = Ek[1][r][lane];
                float x1 = ez_jp - ez, x2 = ey_kp - ey;
                float y1 = ex_kp - ex, y2 = ez_ip - ez;
                float z1 = ey_ip - ey, z2 = ex_jp - ex;
                if (PML) {
                    const float *t = a.tab + a.toff[1];
                    const int sy = slab_index(j, a.J, a.nlo[1], a.nhi[1]);
                    const float byH = t[3 * a.J + j], ayH = t[4 * a.J + j], kyH = t[5 * a.J + j];
                    float tx1 = x1 * kyH, tx2 = x2 * kzH, ty1 = y1 * kzH, ty2 = y2 * kxH, tz1 = z1 * kxH,
                          tz2 = z2 * kyH;
                    if (sx >= 0) {
                        const int q = a.poff[0] + 2 * a.pn[0] + sx + a.W[0] * (j + a.J * kh);
                        ty2 = pml_rw(a, q, bxH, axH, kxH, y2, true);
                        tz1 = pml_rw(a, q + a.pn[0], bxH, axH, kxH, z1, true);
                    }
                    if (sy >= 0) {
                        const int q = a.poff[1] + 2 * a.pn[1] + ic + a.I * (sy + a.W[1] * kh);
                        tz2 = pml_rw(a, q, byH, ayH, kyH, z2, true);
                        tx1 = pml_rw(a, q + a.pn[1], byH, ayH, kyH, x1, true);
                    }
                    if (szh >= 0) {
                        const int q = a.poff[2] + 2 * a.pn[2] + ic + a.I * j + a.P * szh;
                        tx2 = pml_rw(a, q, bzH, azH, kzH, x2, true);
                        ty1 = pml_rw(a, q + a.pn[2], bzH, azH, kzH, y1, true);
                    }
                    x1 = tx1, x2 = tx2, y1 = ty1, y2 = ty2, z1 = tz1, z2 = tz2;
                }
                const float *cf = a.coef + m * kCoefStride;
                float hxn = cf[9] * hxo - cf[12] * (x1 - x2);
                float hyn = cf[10] * hyo - cf[13] * (y1 - y2);
                float hzn = cf[11] * hzo - cf[14] * (z1 - z2);
                float sv[3];
                if (fsource3(a, Mp, kh, c - ph, 3, sv)) {
                    hxn -= cf[15] * sv[0];
                    hyn -= cf[16] * sv[1];
                    hzn -= cf[17] * sv[2];
                }
                D[3 * nal + c] = hxn;
                D[4 * nal + c] = hyn;
                D[5 * nal + c] = hzn;
            }
        }
        bm = bk;
    }
}
}  // namespace

// Copy the current fields into the alternate buffer (boundary walls and all), so that both
// ping-pong buffers agree on every cell the fused kernel does not write.
int fused_sync_alt(fdtd_dev *d) {
    if (!d->alt_base) {
        FDTD_TRY(dev_malloc(&d->alt_base, (size_t)6 * d->Nal * sizeof(float)));
        d->alt_valid = false;
    }
    if (!d->alt_valid) {
        FDTD_TRY(hipMemcpyAsync(d->alt_base, d->field_base, (size_t)6 * d->Nal * sizeof(float),
                                hipMemcpyDeviceToDevice, d->stream));
        d->alt_valid = true;
    }
    return 0;
}

// Second field buffer in sync and second PML memory allocated (zeroed once).
int fused_prepare(fdtd_dev *d) {
    int rc = fused_sync_alt(d);
    if (rc) return rc;
    Cpml &cp = d->cpml;
    const long long pmem = 4 * (cp.psi_n[0] + cp.psi_n[1] + cp.psi_n[2]);
    if (cp.on && cp.psi_base && !cp.psi_alt) {
        FDTD_TRY(dev_malloc(&cp.psi_alt, (size_t)pmem * sizeof(float)));
        FDTD_TRY(hipMemsetAsync(cp.psi_alt, 0, (size_t)pmem * sizeof(float), d->stream));
    }
    return 0;
}

// After a ping-pong step: the second buffers become the current ones.
void fused_swap(fdtd_dev *d) {
    float *t = d->field_base;
    d->field_base = d->alt_base;
    d->alt_base = t;
    for (int m = 0; m < 6; m++) d->F[m] = d->field_base + m * d->Nal;
    Cpml &cp = d->cpml;
    if (cp.on && cp.psi_base && cp.psi_alt) {
        float *old = cp.psi_base;
        cp.psi_base = cp.psi_alt;
        cp.psi_alt = old;
        for (int f = 0; f < 2; f++)
            for (int ax = 0; ax < 3; ax++)
                for (int q = 0; q < 2; q++)
                    if (cp.psi[f][ax][q]) cp.psi[f][ax][q] = cp.psi_base + (cp.psi[f][ax][q] - old);
    }
}

bool fused_supported(const fdtd_dev *d) {
    if (!d->cpml.on || d->dense_src || d->src.size() > 2 || d->mat_bytes > 2) return false;
    const Cpml &cp = d->cpml;
    long long pmem = 4 * (cp.psi_n[0] + cp.psi_n[1] + cp.psi_n[2]);
    return 6 * d->Nal < (1LL << 31) && pmem < (1LL << 31);
}

int fused_step(fdtd_dev *d, double dt, int ty_rows, int kc_len) {
    fields_changed(d);
    int rc = ensure_coefficients(d, dt);
    if (rc) return rc;
    if (d->cpml.on && (rc = cpml_tables(d, dt))) return rc;
    if ((rc = fused_prepare(d))) return rc;
    Cpml &cp = d->cpml;
    FusedArgs a{};
    a.src = d->field_base;
    a.dst = d->alt_base;
    a.nal = (int)d->Nal;
    a.mat = d->mat;
    a.coef = d->d_coef;  // semi-implicit scheme
    a.I = d->I;
    a.J = d->J;
    a.K = d->K;
    a.P = (int)d->P;
    // rows 16 / 32: LDS-tiled kernel with tiles of 63 x 8 / 63 x 16 cells; else register strips
    const bool tiled = ty_rows == 16 || ty_rows == 32;
    const int tile_rows = ty_rows / 2;
    a.nwx = tiled ? (d->I - 2 + kTileOwn - 1) / kTileOwn : (d->I - 2 + kStrip - 1) / kStrip;
    a.nty = tiled ? (d->J - 2 + tile_rows - 1) / tile_rows : (d->J - 2 + ty_rows - 1) / ty_rows;
    a.kc_len = kc_len;
    a.nkc = (d->K - 2 + kc_len - 1) / kc_len;
    a.sk[0] = a.sk[1] = -1;
    a.sJ[0] = a.sJ[1] = nullptr;
    for (size_t s = 0; s < d->src.size() && s < 2; s++) {
        a.sk[s] = d->src[s].k;
        a.sJ[s] = d->src[s].J;
    }
    a.nrs = d->n_run_src;
    for (int q = 0; q < d->n_run_src; q++) a.rs[q] = d->run_src[q];
    const bool pml = cp.on;
    if (pml) {
        a.tab = cp.tab;
        long long off = 0;
        for (int ax = 0; ax < 3; ax++) {
            a.toff[ax] = (int)cp.tab_off[ax];
            a.nlo[ax] = cp.nlo[ax];
            a.nhi[ax] = cp.nhi[ax];
            a.W[ax] = cp.W[ax];
            a.poff[ax] = (int)off;  // same layout as fdtd_dev_cpml_on: axis blocks of 4 arrays
            a.pn[ax] = (int)cp.psi_n[ax];
            off += 4 * cp.psi_n[ax];
        }
        a.psrc = cp.psi_base;
        a.pdst = cp.psi_alt;
    }
    const long long units = (long long)a.nwx * a.nty * a.nkc;  // wavefronts (strips) or blocks (tiles)
    dim3 grid((unsigned)(tiled ? units : (units + 3) / 4));
    const bool u8 = d->mat_bytes == 1;
    hipEvent_t t0, t1;
    timing_events(d, 2, &t0, &t1);
#define FDTD_FUSED_ONE(MT, TYV, PM) \
    hipExtLaunchKernelGGL((k_step_fused<MT, TYV, PM>), grid, dim3(256), 0, d->stream, t0, t1, 0, a)
#define FDTD_FUSED(TYV)                                    \
    do {                                                   \
        if (u8 && pml) FDTD_FUSED_ONE(uint8_t, TYV, true);  \
        else if (u8) FDTD_FUSED_ONE(uint8_t, TYV, false);   \
        else if (pml) FDTD_FUSED_ONE(uint16_t, TYV, true);  \
        else FDTD_FUSED_ONE(uint16_t, TYV, false);          \
    } while (0)
#define FDTD_TILED_ONE(MT, TJV, PM) \
    hipExtLaunchKernelGGL((k_step_tiled<MT, TJV, PM>), grid, dim3(256), 0, d->stream, t0, t1, 0, a)
#define FDTD_TILED(TJV)                                    \
    do {                                                   \
        if (u8 && pml) FDTD_TILED_ONE(uint8_t, TJV, true);  \
        else if (u8) FDTD_TILED_ONE(uint8_t, TJV, false);   \
        else if (pml) FDTD_TILED_ONE(uint16_t, TJV, true);  \
        else FDTD_TILED_ONE(uint16_t, TJV, false);          \
    } while (0)
    switch (ty_rows) {
        case 1: FDTD_FUSED(1); break;
        case 2: FDTD_FUSED(2); break;
        case 4: FDTD_FUSED(4); break;
        case 8: FDTD_FUSED(8); break;
        case 16: FDTD_TILED(8); break;
        case 32: FDTD_TILED(16); break;
        default:
            set_error("fused step: rows must be 1, 2, 4, 8 (register strips) or 16, 32 (LDS tiles)");
            return -2;
    }
#undef FDTD_FUSED
#undef FDTD_FUSED_ONE
#undef FDTD_TILED
#undef FDTD_TILED_ONE
    FDTD_TRY(hipGetLastError());
    fused_swap(d);
    return 0;
}

}  // namespace fdtd
