// Fused E+H step under the PML boundary, one thread per (i, j) column of a z chunk: each field
// component is read once and written once per step (48 B per cell-update instead of the split
// pair's 72 B).  Reference semantics of the update: timefield3d.c:1679-1837 (semi-implicit E/H
// updates); the PML terms are the extension whose formulas oracle/fdtd_oracle.c states.
//
// The H update of cell (i, j, k) needs the new E of its i+1, j+1 and k+1 neighbours.  A thread
// computes them itself from the old fields instead of waiting for their owners: the new Ey, Ez of
// column i+1 and Ex, Ez of row j+1 at each of its planes, and the new Ex, Ey of the plane above
// its chunk.  These redundant values use the same inputs and the same arithmetic as their
// owners', so they are equal to them; the old fields they read are served by L1/L2 (the owners
// read the same lines).  Fields and PML memories are ping-ponged (read the current buffer, write
// the second one), so no thread reads a value another thread has already replaced.  Cells of the
// boundary ring are never written: both buffers hold the same wall values.
//
// As in the split kernels, all loads of a thread's chunk are issued before any is used, there
// is no barrier and no LDS, and the block-to-work map is XCD-aware.
#include "dev_stencil.h"

namespace fdtd {
namespace {

struct ColOut {
    float *D[6];        // second field buffer (written)
    float *psiE[3][2];  // second PML memory (written), same layout as PmlArgs::psiE / psiH
    float *psiH[3][2];
};

struct Tab3 {
    float b, a, k;  // CPML b, a and 1/kappa of one coordinate (0, 0, 1 outside the slab)
};

__device__ __forceinline__ Tab3 tab_at(const float *t, int N, int n, int half, bool in) {
    Tab3 r{0.0f, 0.0f, 1.0f};
    if (in) {
        r.b = t[(3 * half + 0) * N + n];
        r.a = t[(3 * half + 1) * N + n];
        r.k = t[(3 * half + 2) * N + n];
    }
    return r;
}

// PML-corrected difference with the memory read from the current buffer and, when `own`, the
// update written to the second one: psi' = b psi + a d; returns d/kappa + psi' (the split
// kernels' pml_term arithmetic).
template <typename Off>
__device__ __forceinline__ float pml_own(float *dst, Off q, float old, const Tab3 &t, float d, bool own) {
    const float v = t.b * old + t.a * d;
    if (own) st(dst, q, v);
    return d * t.k + v;
}

// The same for a neighbour's memory (recomputed, never stored).
__device__ __forceinline__ float pml_nb(float old, const Tab3 &t, float d) {
    const float v = t.b * old + t.a * d;
    return d * t.k + v;
}

template <typename Off>
__device__ __forceinline__ float ldc(const float *base, Off q, bool in) {
    return in ? ld(base, q) : 0.0f;
}

template <typename MatT, bool PML, bool MON, int KC, typename Off>
__global__ __launch_bounds__(kBlock) void k_step_col(YeeArgs a, PmlArgs pm, ColOut o) {
    const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const int pb = lb % a.plane_blocks, kc = lb / a.plane_blocks;
    const Off p = (Off)pb * (Off)kBlock + (Off)threadIdx.x;
    const int kb = a.k0 + kc * KC;
    const int ke = min(a.k1, kb + KC);
    if (kb >= ke) return;  // uniform over the block
    const int j = (int)(p / (Off)a.I), i = (int)(p - (Off)j * (Off)a.I);
    const bool active = p < (Off)a.P && i >= 1 && i <= a.I - 2 && j >= 1 && j <= a.J - 2;
    if (!MON && !active) return;
    if (active) {
        const float *__restrict__ Ex = a.Ex;
        const float *__restrict__ Ey = a.Ey;
        const float *__restrict__ Ez = a.Ez;
        const float *__restrict__ Hx = a.Hx;
        const float *__restrict__ Hy = a.Hy;
        const float *__restrict__ Hz = a.Hz;
        const Off P = (Off)a.P, I = (Off)a.I;
        const int nv = ke - kb;              // planes of this chunk
        const bool up1 = i + 1 <= a.I - 2;   // column i+1 is updated
        const bool up2 = j + 1 <= a.J - 2;   // row j+1 is updated
        const bool uph = ke < a.k1;          // the plane above the chunk is updated
        int sx = -1, sy = -1, sx1 = -1, sy2 = -1;
        Tab3 xE{0, 0, 1}, yE{0, 0, 1}, xH{0, 0, 1}, yH{0, 0, 1}, x1E{0, 0, 1}, y2E{0, 0, 1};
        if (PML) {
            sx = slab_index(i, a.I, pm.nlo[0], pm.nhi[0]);
            sy = slab_index(j, a.J, pm.nlo[1], pm.nhi[1]);
            sx1 = slab_index(i + 1, a.I, pm.nlo[0], pm.nhi[0]);
            sy2 = slab_index(j + 1, a.J, pm.nlo[1], pm.nhi[1]);
            xE = tab_at(pm.tab[0], a.I, i, 0, sx >= 0);
            xH = tab_at(pm.tab[0], a.I, i, 1, sx >= 0);
            yE = tab_at(pm.tab[1], a.J, j, 0, sy >= 0);
            yH = tab_at(pm.tab[1], a.J, j, 1, sy >= 0);
            x1E = tab_at(pm.tab[0], a.I, i + 1, 0, sx1 >= 0);
            y2E = tab_at(pm.tab[1], a.J, j + 1, 0, sy2 >= 0);
        }

        // ---- loads: every plane of the chunk, its i+1 column and j+1 row, and the plane above ----
        const Off c0 = p + (Off)kb * P;
        const float hx_k0 = ld(Hx, c0 - P), hy_k0 = ld(Hy, c0 - P);
        const float hx1_k0 = ld(Hx, c0 + 1 - P), hy2_k0 = ld(Hy, c0 + I - P);
        float eox[KC], eoy[KC], eoz[KC], hx[KC], hy[KC], hz[KC], hy_im[KC], hz_im[KC], hx_jm[KC], hz_jm[KC];
        float eoy1[KC], eoz1[KC], hx1[KC], hy1[KC], hz1[KC], hx1_jm[KC];
        float eox2[KC], eoz2[KC], hx2[KC], hy2[KC], hz2[KC], hy2_im[KC];
        int m[KC], m1[KC], m2[KC];
        float pe[KC][6], ph[KC][6], pn1[KC][4], pn2[KC][4];
#pragma unroll
        for (int t = 0; t < KC; t++) {
            const int kk = t < nv ? kb + t : ke - 1;
            const Off c = p + (Off)kk * P;
            eox[t] = ld(Ex, c), eoy[t] = ld(Ey, c), eoz[t] = ld(Ez, c);
            hx[t] = ld(Hx, c), hy[t] = ld(Hy, c), hz[t] = ld(Hz, c);
            hy_im[t] = ld(Hy, c - 1), hz_im[t] = ld(Hz, c - 1);
            hx_jm[t] = ld(Hx, c - I), hz_jm[t] = ld(Hz, c - I);
            m[t] = mat_at<MatT>(a.mat, c);
            eoy1[t] = ld(Ey, c + 1), eoz1[t] = ld(Ez, c + 1);
            hx1[t] = ld(Hx, c + 1), hy1[t] = ld(Hy, c + 1), hz1[t] = ld(Hz, c + 1);
            hx1_jm[t] = ld(Hx, c + 1 - I);
            m1[t] = mat_at<MatT>(a.mat, c + 1);
            eox2[t] = ld(Ex, c + I), eoz2[t] = ld(Ez, c + I);
            hx2[t] = ld(Hx, c + I), hy2[t] = ld(Hy, c + I), hz2[t] = ld(Hz, c + I);
            hy2_im[t] = ld(Hy, c + I - 1);
            m2[t] = mat_at<MatT>(a.mat, c + I);
            if (PML) {
                psi_load(pm.psiE, pm, a, i, j, p, kk, sx, sy, pe[t]);
                psi_load(pm.psiH, pm, a, i, j, p, kk, sx, sy, ph[t]);
                const int sz = slab_index(kk, a.K, pm.nlo[2], pm.nhi[2]);
                const Off qx1 = psi_x<Off>(pm, a, j, kk, sx1), qy1 = psi_y<Off>(pm, a, i + 1, kk, sy);
                const Off qx2 = psi_x<Off>(pm, a, j + 1, kk, sx), qy2 = psi_y<Off>(pm, a, i, kk, sy2);
                pn1[t][0] = ldc(pm.psiE[0][0], qx1, up1 && sx1 >= 0);  // Ey, x term
                pn1[t][1] = ldc(pm.psiE[0][1], qx1, up1 && sx1 >= 0);  // Ez, x term
                pn1[t][2] = ldc(pm.psiE[1][0], qy1, up1 && sy >= 0);   // Ez, y term
                pn1[t][3] = ldc(pm.psiE[2][1], psi_z<Off>(a, p + 1, sz), up1 && sz >= 0);  // Ey, z term
                pn2[t][0] = ldc(pm.psiE[0][1], qx2, up2 && sx >= 0);   // Ez, x term
                pn2[t][1] = ldc(pm.psiE[1][0], qy2, up2 && sy2 >= 0);  // Ez, y term
                pn2[t][2] = ldc(pm.psiE[1][1], qy2, up2 && sy2 >= 0);  // Ex, y term
                pn2[t][3] = ldc(pm.psiE[2][0], psi_z<Off>(a, p + I, sz), up2 && sz >= 0);  // Ex, z term
            }
        }
        // the plane above the chunk: its Ex, Ey (new when updated, else the wall value)
        const Off chh = p + (Off)ke * P;
        const float eox_h = ld(Ex, chh), eoy_h = ld(Ey, chh);
        float hx_h = 0, hy_h = 0, hz_h = 0, hz_im_h = 0, hz_jm_h = 0, php[4] = {0, 0, 0, 0};
        int m_h = 0, szh = -1;
        if (uph) {
            hx_h = ld(Hx, chh), hy_h = ld(Hy, chh), hz_h = ld(Hz, chh);
            hz_im_h = ld(Hz, chh - 1), hz_jm_h = ld(Hz, chh - I);
            m_h = mat_at<MatT>(a.mat, chh);
            if (PML) {
                szh = slab_index(ke, a.K, pm.nlo[2], pm.nhi[2]);
                php[0] = ldc(pm.psiE[0][0], psi_x<Off>(pm, a, j, ke, sx), sx >= 0);       // Ey, x term
                php[1] = ldc(pm.psiE[1][1], psi_y<Off>(pm, a, i, ke, sy), sy >= 0);       // Ex, y term
                php[2] = ldc(pm.psiE[2][0], psi_z<Off>(a, p, szh), szh >= 0);             // Ex, z term
                php[3] = ldc(pm.psiE[2][1], psi_z<Off>(a, p, szh), szh >= 0);             // Ey, z term
            }
        }

        // ---- new E: own column, column i+1 (Ey, Ez), row j+1 (Ex, Ez) ----
        float ex[KC], ey[KC], ez[KC], ey1[KC], ez1[KC], ex2[KC], ez2[KC];
#pragma unroll
        for (int t = 0; t < KC; t++) {
            const bool live = t < nv;
            const int k = live ? kb + t : ke - 1;
            const Off c = p + (Off)k * P;
            int sz = -1;
            Tab3 zE{0, 0, 1};
            if (PML) {
                sz = slab_index(k, a.K, pm.nlo[2], pm.nhi[2]);
                zE = tab_at(pm.tab[2], a.K, k, 0, sz >= 0);
            }
            const float hxk = t == 0 ? hx_k0 : hx[t - 1], hyk = t == 0 ? hy_k0 : hy[t - 1];
            float sv[3];
            {  // own cell (update_E_block's arithmetic)
                float x1 = hz[t] - hz_jm[t], x2 = hy[t] - hyk;
                float y1 = hx[t] - hxk, y2 = hz[t] - hz_im[t];
                float z1 = hy[t] - hy_im[t], z2 = hx[t] - hx_jm[t];
                if (PML && (sx >= 0 || sy >= 0 || sz >= 0)) {
                    float tx1 = x1 * yE.k, tx2 = x2 * zE.k, ty1 = y1 * zE.k, ty2 = y2 * xE.k, tz1 = z1 * xE.k,
                          tz2 = z2 * yE.k;
                    if (sx >= 0) {
                        const Off q = psi_x<Off>(pm, a, j, k, sx);
                        ty2 = pml_own(o.psiE[0][0], q, pe[t][0], xE, y2, live);
                        tz1 = pml_own(o.psiE[0][1], q, pe[t][1], xE, z1, live);
                    }
                    if (sy >= 0) {
                        const Off q = psi_y<Off>(pm, a, i, k, sy);
                        tz2 = pml_own(o.psiE[1][0], q, pe[t][2], yE, z2, live);
                        tx1 = pml_own(o.psiE[1][1], q, pe[t][3], yE, x1, live);
                    }
                    if (sz >= 0) {
                        const Off q = psi_z<Off>(a, p, sz);
                        tx2 = pml_own(o.psiE[2][0], q, pe[t][4], zE, x2, live);
                        ty1 = pml_own(o.psiE[2][1], q, pe[t][5], zE, y1, live);
                    }
                    x1 = tx1, x2 = tx2, y1 = ty1, y2 = ty2, z1 = tz1, z2 = tz2;
                }
                const float *cf = a.coef + m[t] * kCoefStride;
                ex[t] = cf[0] * eox[t] + cf[3] * (x1 - x2);
                ey[t] = cf[1] * eoy[t] + cf[4] * (y1 - y2);
                ez[t] = cf[2] * eoz[t] + cf[5] * (z1 - z2);
                if (source3(a, k, p, 0, sv)) {
                    ex[t] -= cf[6] * sv[0];
                    ey[t] -= cf[7] * sv[1];
                    ez[t] -= cf[8] * sv[2];
                }
                if (live) {
                    st(o.D[EX], c, ex[t]);
                    st(o.D[EY], c, ey[t]);
                    st(o.D[EZ], c, ez[t]);
                }
            }
            ey1[t] = eoy1[t], ez1[t] = eoz1[t];
            if (up1) {  // column i+1: Ey, Ez
                const float hx1k = t == 0 ? hx1_k0 : hx1[t - 1];
                float y1 = hx1[t] - hx1k, y2 = hz1[t] - hz[t];
                float z1 = hy1[t] - hy[t], z2 = hx1[t] - hx1_jm[t];
                if (PML && (sx1 >= 0 || sy >= 0 || sz >= 0)) {
                    float ty1 = y1 * zE.k, ty2 = y2 * x1E.k, tz1 = z1 * x1E.k, tz2 = z2 * yE.k;
                    if (sx1 >= 0) {
                        ty2 = pml_nb(pn1[t][0], x1E, y2);
                        tz1 = pml_nb(pn1[t][1], x1E, z1);
                    }
                    if (sy >= 0) tz2 = pml_nb(pn1[t][2], yE, z2);
                    if (sz >= 0) ty1 = pml_nb(pn1[t][3], zE, y1);
                    y1 = ty1, y2 = ty2, z1 = tz1, z2 = tz2;
                }
                const float *cf = a.coef + m1[t] * kCoefStride;
                ey1[t] = cf[1] * eoy1[t] + cf[4] * (y1 - y2);
                ez1[t] = cf[2] * eoz1[t] + cf[5] * (z1 - z2);
                if (source3(a, k, p + 1, 0, sv)) {
                    ey1[t] -= cf[7] * sv[1];
                    ez1[t] -= cf[8] * sv[2];
                }
            }
            ex2[t] = eox2[t], ez2[t] = eoz2[t];
            if (up2) {  // row j+1: Ex, Ez
                const float hy2k = t == 0 ? hy2_k0 : hy2[t - 1];
                float x1 = hz2[t] - hz[t], x2 = hy2[t] - hy2k;
                float z1 = hy2[t] - hy2_im[t], z2 = hx2[t] - hx[t];
                if (PML && (sx >= 0 || sy2 >= 0 || sz >= 0)) {
                    float tx1 = x1 * y2E.k, tx2 = x2 * zE.k, tz1 = z1 * xE.k, tz2 = z2 * y2E.k;
                    if (sx >= 0) tz1 = pml_nb(pn2[t][0], xE, z1);
                    if (sy2 >= 0) {
                        tz2 = pml_nb(pn2[t][1], y2E, z2);
                        tx1 = pml_nb(pn2[t][2], y2E, x1);
                    }
                    if (sz >= 0) tx2 = pml_nb(pn2[t][3], zE, x2);
                    x1 = tx1, x2 = tx2, z1 = tz1, z2 = tz2;
                }
                const float *cf = a.coef + m2[t] * kCoefStride;
                ex2[t] = cf[0] * eox2[t] + cf[3] * (x1 - x2);
                ez2[t] = cf[2] * eoz2[t] + cf[5] * (z1 - z2);
                if (source3(a, k, p + I, 0, sv)) {
                    ex2[t] -= cf[6] * sv[0];
                    ez2[t] -= cf[8] * sv[2];
                }
            }
        }
        float exh = eox_h, eyh = eoy_h;
        if (uph) {  // the plane above: Ex, Ey
            float hxl = hx[0], hyl = hy[0];
#pragma unroll
            for (int t = 1; t < KC; t++)
                if (t < nv) hxl = hx[t], hyl = hy[t];
            Tab3 zE{0, 0, 1};
            if (PML) zE = tab_at(pm.tab[2], a.K, ke, 0, szh >= 0);
            float x1 = hz_h - hz_jm_h, x2 = hy_h - hyl;
            float y1 = hx_h - hxl, y2 = hz_h - hz_im_h;
            if (PML && (sx >= 0 || sy >= 0 || szh >= 0)) {
                float tx1 = x1 * yE.k, tx2 = x2 * zE.k, ty1 = y1 * zE.k, ty2 = y2 * xE.k;
                if (sx >= 0) ty2 = pml_nb(php[0], xE, y2);
                if (sy >= 0) tx1 = pml_nb(php[1], yE, x1);
                if (szh >= 0) {
                    tx2 = pml_nb(php[2], zE, x2);
                    ty1 = pml_nb(php[3], zE, y1);
                }
                x1 = tx1, x2 = tx2, y1 = ty1, y2 = ty2;
            }
            const float *cf = a.coef + m_h * kCoefStride;
            exh = cf[0] * eox_h + cf[3] * (x1 - x2);
            eyh = cf[1] * eoy_h + cf[4] * (y1 - y2);
            float sv[3];
            if (source3(a, ke, p, 0, sv)) {
                exh -= cf[6] * sv[0];
                eyh -= cf[7] * sv[1];
            }
        }

        // ---- new H of the chunk (update_H_block's arithmetic) ----
#pragma unroll
        for (int t = 0; t < KC; t++) {
            const bool live = t < nv;
            const int k = live ? kb + t : ke - 1;
            const Off c = p + (Off)k * P;
            const float exk = (t + 1 < KC && t + 1 < nv) ? ex[t + 1 < KC ? t + 1 : t] : exh;
            const float eyk = (t + 1 < KC && t + 1 < nv) ? ey[t + 1 < KC ? t + 1 : t] : eyh;
            float x1 = ez2[t] - ez[t], x2 = eyk - ey[t];
            float y1 = exk - ex[t], y2 = ez1[t] - ez[t];
            float z1 = ey1[t] - ey[t], z2 = ex2[t] - ex[t];
            const int sz = PML ? slab_index(k, a.K, pm.nlo[2], pm.nhi[2]) : -1;
            if (PML && (sx >= 0 || sy >= 0 || sz >= 0)) {
                const Tab3 zH = tab_at(pm.tab[2], a.K, k, 1, sz >= 0);
                float tx1 = x1 * yH.k, tx2 = x2 * zH.k, ty1 = y1 * zH.k, ty2 = y2 * xH.k, tz1 = z1 * xH.k,
                      tz2 = z2 * yH.k;
                if (sx >= 0) {
                    const Off q = psi_x<Off>(pm, a, j, k, sx);
                    ty2 = pml_own(o.psiH[0][0], q, ph[t][0], xH, y2, live);
                    tz1 = pml_own(o.psiH[0][1], q, ph[t][1], xH, z1, live);
                }
                if (sy >= 0) {
                    const Off q = psi_y<Off>(pm, a, i, k, sy);
                    tz2 = pml_own(o.psiH[1][0], q, ph[t][2], yH, z2, live);
                    tx1 = pml_own(o.psiH[1][1], q, ph[t][3], yH, x1, live);
                }
                if (sz >= 0) {
                    const Off q = psi_z<Off>(a, p, sz);
                    tx2 = pml_own(o.psiH[2][0], q, ph[t][4], zH, x2, live);
                    ty1 = pml_own(o.psiH[2][1], q, ph[t][5], zH, y1, live);
                }
                x1 = tx1, x2 = tx2, y1 = ty1, y2 = ty2, z1 = tz1, z2 = tz2;
            }
            const float *cf = a.coef + m[t] * kCoefStride;
            float hxn = cf[9] * hx[t] - cf[12] * (x1 - x2);
            float hyn = cf[10] * hy[t] - cf[13] * (y1 - y2);
            float hzn = cf[11] * hz[t] - cf[14] * (z1 - z2);
            float sv[3];
            if (source3(a, k, p, 3, sv)) {
                hxn -= cf[15] * sv[0];
                hyn -= cf[16] * sv[1];
                hzn -= cf[17] * sv[2];
            }
            if (live) {
                st(o.D[HX], c, hxn);
                st(o.D[HY], c, hyn);
                st(o.D[HZ], c, hzn);
            }
        }
    }
    if (MON) monitor_epilogue_on(a, o.D[EX], o.D[EY], o.D[HX], o.D[HY], kb, ke, pb, p, active);
}

template <int KC>
void launch_col(bool mon, bool u8, dim3 grid, hipStream_t st, hipEvent_t t0, hipEvent_t t1, const YeeArgs &a,
                const PmlArgs &pm, const ColOut &o) {
#define FDTD_GO(KER) hipExtLaunchKernelGGL(KER, grid, dim3(kBlock), 0, st, t0, t1, 0, a, pm, o)
    if (mon) {
        if (u8) FDTD_GO((k_step_col<uint8_t, true, true, KC, uint32_t>));
        else FDTD_GO((k_step_col<uint16_t, true, true, KC, uint32_t>));
    } else {
        if (u8) FDTD_GO((k_step_col<uint8_t, true, false, KC, uint32_t>));
        else FDTD_GO((k_step_col<uint16_t, true, false, KC, uint32_t>));
    }
#undef FDTD_GO
}

}  // namespace

bool col_supported(const fdtd_dev *d) {
    return d->cpml.on && d->mat_bytes <= 2 && d->N < kSmallGridCells && d->K >= 3;
}

int col_step(fdtd_dev *d, double dt, int kc) {
    fields_changed(d);
    if (kc != 1 && kc != 2 && kc != 4) {
        set_error("column step: z planes per thread must be 1, 2 or 4");
        return -2;
    }
    YeeArgs a;
    PmlArgs pm;
    int rc = fill_step_args(d, dt, &a, &pm);
    if (rc) return rc;
    if ((rc = fused_prepare(d))) return rc;
    a.kchunk = kc;
    a.nrs = d->n_run_src;
    for (int q = 0; q < d->n_run_src; q++) a.rs[q] = d->run_src[q];
    set_source_range(a);
    const bool mon = d->n_run_mon > 0;
    a.nmon = mon ? d->n_run_mon : 0;
    for (int q = 0; q < a.nmon; q++) a.mon[q] = d->run_mon[q];
    if (mon) d->n_run_mon = 0;  // consumed
    const Cpml &cp = d->cpml;
    ColOut o{};
    for (int m = 0; m < 6; m++) o.D[m] = d->alt_base + m * d->Nal;
    for (int ax = 0; ax < 3; ax++)
        for (int q = 0; q < 2; q++) {
            o.psiE[ax][q] = cp.psi[0][ax][q] ? cp.psi_alt + (cp.psi[0][ax][q] - cp.psi_base) : nullptr;
            o.psiH[ax][q] = cp.psi[1][ax][q] ? cp.psi_alt + (cp.psi[1][ax][q] - cp.psi_base) : nullptr;
        }
    const int nz = a.k1 - a.k0;
    dim3 grid((unsigned)(a.plane_blocks * ((nz + kc - 1) / kc)));
    const bool u8 = d->mat_bytes == 1;
    hipEvent_t t0, t1;
    timing_events(d, 2, &t0, &t1);
    switch (kc) {
        case 1: launch_col<1>(mon, u8, grid, d->stream, t0, t1, a, pm, o); break;
        case 2: launch_col<2>(mon, u8, grid, d->stream, t0, t1, a, pm, o); break;
        default: launch_col<4>(mon, u8, grid, d->stream, t0, t1, a, pm, o); break;
    }
    FDTD_TRY(hipGetLastError());
    fused_swap(d);
    return 0;
}

}  // namespace fdtd
