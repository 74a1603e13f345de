// Yee leapfrog on gfx950: split E / H z-marching kernels (optionally with the convolutional PML),
// boundary-shell snapshot and the first-order Mur faces.  Reference semantics:
// timefield3d.c:246-346 (shell snapshot), 1537-1837 (E/H updates), 2275-2721 (Mur faces),
// 2723-2849 (step drivers).  The PML is an extension with no reference counterpart; its
// formulas are the ones stated in oracle/fdtd_oracle.c (Roden-Gedney CPML with CFS grading).
//
// Layout: six FP32 component arrays, cell c = i + I*(j + J*k).  A thread owns one (i, j) column of
// a z chunk and marches k upward, keeping the k-1 (E update) or k+1 (H update) neighbours it
// needs in registers; the i-1/j-1 (i+1/j+1) neighbours come through L1/L2.  Per cell-update the
// E pass streams E in/out, H in and a 1-byte material id: 25 bytes of HBM traffic.
#include <cmath>

#include "dev_internal.h"

namespace fdtd {
namespace {

struct YeeArgs {
    float *Ex, *Ey, *Ez, *Hx, *Hy, *Hz;
    const void *mat;
    const float *coef;  // [nmat][kCoefStride] for the chosen scheme
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
};

// Source components m0..m0+2 of plane cell p at plane k: tabulated run source inside its mode
// window, else the stored source plane.  Returns false when plane k carries no source.
__device__ __forceinline__ bool source3(const YeeArgs &a, int k, long long p, int m0, float v[3]);

struct PmlArgs {
    const float *tab[3];  // per axis: bE, aE, kE, bH, aH, kH (N_axis floats each)
    int nlo[3], nhi[3], W[3];
    float *psiE[3][2];
    float *psiH[3][2];
};

template <typename MatT>
__device__ __forceinline__ int mat_id(const void *mat, long long c) {
    return (int)static_cast<const MatT *>(mat)[c];
}

// Source values of plane k (component m of plane cell p at J[m*stride + p]), or null.
__device__ __forceinline__ const float *src_plane(const YeeArgs &a, int k, long long &stride) {
    if (a.dense) {
        stride = a.N;
        return a.dense + (long long)k * a.P;
    }
    const float *J = nullptr;
    stride = a.P;
    for (int s = 0; s < a.nsrc; s++)
        if (a.src[s].k == k) J = a.src[s].J;
    return J;
}

__device__ __forceinline__ bool source3(const YeeArgs &a, int k, long long p, int m0, float v[3]) {
    for (int q = 0; q < a.nrs; q++) {
        const RunSrcArg &r = a.rs[q];
        if (r.k == k && r.win[p]) {
            for (int m = 0; m < 3; m++) v[m] = r.c1 * r.PR[(m0 + m) * a.P + p] + r.c2 * r.PI[(m0 + m) * a.P + p];
            return true;
        }
    }
    long long cs;
    const float *J = src_plane(a, k, cs);
    if (!J) return false;
    for (int m = 0; m < 3; m++) v[m] = J[(m0 + m) * cs + p];
    return true;
}

__device__ __forceinline__ int slab_index(int n, int N, int nlo, int nhi) {
    if (n < nlo) return n;
    if (nhi > 0 && n >= N - 1 - nhi) return nlo + n - (N - 1 - nhi);
    return -1;
}

// psi <- b psi + a d; returns the PML-corrected difference d/kappa + psi
__device__ __forceinline__ float pml_term(float *psi, float b, float a, float kinv, float d) {
    float v = b * *psi + a * d;
    *psi = v;
    return d * kinv + v;
}

// E <- caE*E + cbE*(backward-difference curl of H) - cjE*J   (timefield3d.c:1759-1837)
template <typename MatT, bool PML>
__global__ __launch_bounds__(kBlock) void k_update_E(YeeArgs a, PmlArgs pm) {
    const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const int pb = lb % a.plane_blocks, kc = lb / a.plane_blocks;
    const long long p = (long long)pb * kBlock + threadIdx.x;
    if (p >= a.P) return;
    const int j = (int)(p / a.I), i = (int)(p - (long long)j * a.I);
    if (i < 1 || i > a.I - 2 || j < 1 || j > a.J - 2) return;
    const int kb = a.k0 + kc * a.kchunk;
    const int ke = min(a.k1, kb + a.kchunk);
    if (kb >= ke) return;
    int sx = -1, sy = -1;
    float bx = 0, ax = 0, kx = 1, by = 0, ay = 0, ky = 1;
    if (PML) {
        sx = slab_index(i, a.I, pm.nlo[0], pm.nhi[0]);
        sy = slab_index(j, a.J, pm.nlo[1], pm.nhi[1]);
        bx = pm.tab[0][i], ax = pm.tab[0][a.I + i], kx = pm.tab[0][2 * a.I + i];
        by = pm.tab[1][j], ay = pm.tab[1][a.J + j], ky = pm.tab[1][2 * a.J + j];
    }
    long long c = p + (long long)kb * a.P;
    float hx_km = a.Hx[c - a.P], hy_km = a.Hy[c - a.P];
    for (int k = kb; k < ke; k++, c += a.P) {
        const float hx = a.Hx[c], hy = a.Hy[c], hz = a.Hz[c];
        const float hy_im = a.Hy[c - 1], hz_im = a.Hz[c - 1];
        const float hx_jm = a.Hx[c - a.I], hz_jm = a.Hz[c - a.I];
        const float *cf = a.coef + (long long)mat_id<MatT>(a.mat, c) * kCoefStride;
        // curl_x = dHz/dy - dHy/dz, curl_y = dHx/dz - dHz/dx, curl_z = dHy/dx - dHx/dy
        float x1 = hz - hz_jm, x2 = hy - hy_km;
        float y1 = hx - hx_km, y2 = hz - hz_im;
        float z1 = hy - hy_im, z2 = hx - hx_jm;
        if (PML) {
            const int sz = slab_index(k, a.K, pm.nlo[2], pm.nhi[2]);
            const float kz = pm.tab[2][2 * a.K + k];
            float tx1 = x1 * ky, tx2 = x2 * kz, ty1 = y1 * kz, ty2 = y2 * kx, tz1 = z1 * kx, tz2 = z2 * ky;
            if (sx >= 0) {
                const long long q = sx + (long long)pm.W[0] * (j + (long long)a.J * k);
                ty2 = pml_term(pm.psiE[0][0] + q, bx, ax, kx, y2);
                tz1 = pml_term(pm.psiE[0][1] + q, bx, ax, kx, z1);
            }
            if (sy >= 0) {
                const long long q = i + (long long)a.I * (sy + (long long)pm.W[1] * k);
                tz2 = pml_term(pm.psiE[1][0] + q, by, ay, ky, z2);
                tx1 = pml_term(pm.psiE[1][1] + q, by, ay, ky, x1);
            }
            if (sz >= 0) {
                const float bz = pm.tab[2][k], az = pm.tab[2][a.K + k];
                const long long q = p + a.P * sz;
                tx2 = pml_term(pm.psiE[2][0] + q, bz, az, kz, x2);
                ty1 = pml_term(pm.psiE[2][1] + q, bz, az, kz, y1);
            }
            x1 = tx1, x2 = tx2, y1 = ty1, y2 = ty2, z1 = tz1, z2 = tz2;
        }
        float ex = cf[0] * a.Ex[c] + cf[3] * (x1 - x2);
        float ey = cf[1] * a.Ey[c] + cf[4] * (y1 - y2);
        float ez = cf[2] * a.Ez[c] + cf[5] * (z1 - z2);
        float sv[3];
        if (source3(a, k, p, 0, sv)) {
            ex -= cf[6] * sv[0];
            ey -= cf[7] * sv[1];
            ez -= cf[8] * sv[2];
        }
        a.Ex[c] = ex;
        a.Ey[c] = ey;
        a.Ez[c] = ez;
        hx_km = hx;
        hy_km = hy;
    }
}

// H <- caH*H - cbH*(forward-difference curl of E) - cjH*M   (timefield3d.c:1679-1757)
template <typename MatT, bool PML>
__global__ __launch_bounds__(kBlock) void k_update_H(YeeArgs a, PmlArgs pm) {
    const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const int pb = lb % a.plane_blocks, kc = lb / a.plane_blocks;
    const long long p = (long long)pb * kBlock + threadIdx.x;
    if (p >= a.P) return;
    const int j = (int)(p / a.I), i = (int)(p - (long long)j * a.I);
    if (i < 1 || i > a.I - 2 || j < 1 || j > a.J - 2) return;
    const int kb = a.k0 + kc * a.kchunk;
    const int ke = min(a.k1, kb + a.kchunk);
    if (kb >= ke) return;
    int sx = -1, sy = -1;
    float bx = 0, ax = 0, kx = 1, by = 0, ay = 0, ky = 1;
    if (PML) {
        sx = slab_index(i, a.I, pm.nlo[0], pm.nhi[0]);
        sy = slab_index(j, a.J, pm.nlo[1], pm.nhi[1]);
        bx = pm.tab[0][3 * a.I + i], ax = pm.tab[0][4 * a.I + i], kx = pm.tab[0][5 * a.I + i];
        by = pm.tab[1][3 * a.J + j], ay = pm.tab[1][4 * a.J + j], ky = pm.tab[1][5 * a.J + j];
    }
    long long c = p + (long long)kb * a.P;
    float ex = a.Ex[c], ey = a.Ey[c];
    for (int k = kb; k < ke; k++, c += a.P) {
        const float ez = a.Ez[c];
        const float ex_kp = a.Ex[c + a.P], ey_kp = a.Ey[c + a.P];
        const float ey_ip = a.Ey[c + 1], ez_ip = a.Ez[c + 1];
        const float ex_jp = a.Ex[c + a.I], ez_jp = a.Ez[c + a.I];
        const float *cf = a.coef + (long long)mat_id<MatT>(a.mat, c) * kCoefStride;
        // curl_x = dEz/dy - dEy/dz, curl_y = dEx/dz - dEz/dx, curl_z = dEy/dx - dEx/dy
        float x1 = ez_jp - ez, x2 = ey_kp - ey;
        float y1 = ex_kp - ex, y2 = ez_ip - ez;
        float z1 = ey_ip - ey, z2 = ex_jp - ex;
        if (PML) {
            const int sz = slab_index(k, a.K, pm.nlo[2], pm.nhi[2]);
            const float kz = pm.tab[2][5 * a.K + k];
            float tx1 = x1 * ky, tx2 = x2 * kz, ty1 = y1 * kz, ty2 = y2 * kx, tz1 = z1 * kx, tz2 = z2 * ky;
            if (sx >= 0) {
                const long long q = sx + (long long)pm.W[0] * (j + (long long)a.J * k);
                ty2 = pml_term(pm.psiH[0][0] + q, bx, ax, kx, y2);
                tz1 = pml_term(pm.psiH[0][1] + q, bx, ax, kx, z1);
            }
            if (sy >= 0) {
                const long long q = i + (long long)a.I * (sy + (long long)pm.W[1] * k);
                tz2 = pml_term(pm.psiH[1][0] + q, by, ay, ky, z2);
                tx1 = pml_term(pm.psiH[1][1] + q, by, ay, ky, x1);
            }
            if (sz >= 0) {
                const float bz = pm.tab[2][3 * a.K + k], az = pm.tab[2][4 * a.K + k];
                const long long q = p + a.P * sz;
                tx2 = pml_term(pm.psiH[2][0] + q, bz, az, kz, x2);
                ty1 = pml_term(pm.psiH[2][1] + q, bz, az, kz, y1);
            }
            x1 = tx1, x2 = tx2, y1 = ty1, y2 = ty2, z1 = tz1, z2 = tz2;
        }
        float hx = cf[9] * a.Hx[c] - cf[12] * (x1 - x2);
        float hy = cf[10] * a.Hy[c] - cf[13] * (y1 - y2);
        float hz = cf[11] * a.Hz[c] - cf[14] * (z1 - z2);
        float sv[3];
        if (source3(a, k, p, 3, sv)) {
            hx -= cf[15] * sv[0];
            hy -= cf[16] * sv[1];
            hz -= cf[17] * sv[2];
        }
        a.Hx[c] = hx;
        a.Hy[c] = hy;
        a.Hz[c] = hz;
        ex = ex_kp;
        ey = ey_kp;
    }
}

struct FaceArgs {
    float *F[6];
    float *shell;
    long long shell_off[6];
    int N[3];
    const void *mat;
    int mat_bytes;
    const float *coef;
};

__device__ __forceinline__ long long cell_of(const int x[3], const int N[3]) {
    return (long long)x[0] + (long long)N[0] * ((long long)x[1] + (long long)N[1] * x[2]);
}

// Copy the tangential components of the two outermost layers of every face into the shell
// buffer (timefield3d.c:246-346 copies all six; only the tangential pairs are ever read back).
__global__ void k_snapshot(FaceArgs fa, long long total, long long a0, long long a1, long long a2) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const long long area[3] = {a0, a1, a2};
    int f = 0;
    long long r = t;
    while (r >= 2 * area[f % 3]) {
        r -= 2 * area[f % 3];
        f++;
    }
    const int a = f % 3, side = f / 3, b = (a + 1) % 3, cc = (a + 2) % 3;
    const int layer = (int)(r / area[a]);
    const long long uv = r - layer * area[a];
    const int u = (int)(uv % fa.N[b]), v = (int)(uv / fa.N[b]);
    int x[3];
    x[a] = side == 0 ? layer : fa.N[a] - 1 - layer;
    x[b] = u;
    x[cc] = v;
    const long long c = cell_of(x, fa.N);
    float *sh = fa.shell + fa.shell_off[f] + (long long)layer * 4 * area[a] + uv;
    sh[0 * area[a]] = fa.F[b][c];
    sh[1 * area[a]] = fa.F[cc][c];
    sh[2 * area[a]] = fa.F[3 + b][c];
    sh[3 * area[a]] = fa.F[3 + cc][c];
}

// One Mur face (axis a, side), field 0 = E / 1 = H, restricted to z in [k0, k1) for x and y
// faces: F_b <- F_in^old + s (F_in - F_b^old)   (timefield3d.c:2275-2721).
template <typename MatT>
__global__ void k_mur_face(FaceArgs fa, int a, int side, int field, int k0, int k1) {
    const int b = (a + 1) % 3, cc = (a + 2) % 3;
    const long long area = (long long)fa.N[b] * fa.N[cc];
    // enumerate only the z-restricted part of the face
    int ulo = 0, uhi = fa.N[b], vlo = 0, vhi = fa.N[cc];
    if (b == 2) { ulo = k0; uhi = k1; }
    if (cc == 2) { vlo = k0; vhi = k1; }
    const long long nu = uhi - ulo;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nu * (vhi - vlo)) return;
    const int u = ulo + (int)(t % nu), v = vlo + (int)(t / nu);
    int xb[3], xi[3];
    xb[a] = side == 0 ? 0 : fa.N[a] - 1;
    xi[a] = side == 0 ? 1 : fa.N[a] - 2;
    xb[b] = xi[b] = u;
    xb[cc] = xi[cc] = v;
    const long long cb = cell_of(xb, fa.N), ci = cell_of(xi, fa.N);
    const long long uv = u + (long long)fa.N[b] * v;
    const int f = a + 3 * side;
    const float *sh0 = fa.shell + fa.shell_off[f] + uv;          // boundary layer
    const float *sh1 = sh0 + 4 * area;                          // inner layer
    const float *cf = fa.coef + (long long)mat_id<MatT>(fa.mat, cb) * kCoefStride + 18 + 6 * field + 2 * a;
    const int comp0 = 3 * field;
    // q = 0 -> component b, q = 1 -> component cc
    {
        float *Fq = fa.F[comp0 + b];
        Fq[cb] = sh1[(2 * field + 0) * area] + cf[0] * (Fq[ci] - sh0[(2 * field + 0) * area]);
    }
    {
        float *Fq = fa.F[comp0 + cc];
        Fq[cb] = sh1[(2 * field + 1) * area] + cf[1] * (Fq[ci] - sh0[(2 * field + 1) * area]);
    }
}

FaceArgs face_args(fdtd_dev *d, int scheme) {
    FaceArgs fa;
    for (int m = 0; m < 6; m++) fa.F[m] = d->F[m];
    fa.shell = d->shell;
    for (int f = 0; f < 6; f++) fa.shell_off[f] = d->shell_off[f];
    fa.N[0] = d->I;
    fa.N[1] = d->J;
    fa.N[2] = d->K;
    fa.mat = d->mat;
    fa.mat_bytes = d->mat_bytes;
    fa.coef = d->d_coef + (size_t)scheme * d->nmat * kCoefStride;
    return fa;
}

// Host restatement of the CPML grading (same formulas as oracle/fdtd_oracle.c:orc_cpml_profile),
// evaluated in FP64 and rounded once to FP32.
void cpml_profile(int N, int nlo, int nhi, const Cpml &c, double disc, double dt, float *out) {
    for (int n = 0; n < N; n++) {
        for (int half = 0; half < 2; half++) {
            double x = n + 0.5 * half;
            bool in_slab = half == 0 ? (n < nlo || (nhi > 0 && n >= N - nhi))
                                     : (n < nlo || (nhi > 0 && n >= N - 1 - nhi));
            double b = 0.0, a = 0.0, kinv = 1.0;
            if (in_slab) {
                int npml = n < nlo ? nlo : nhi;
                double rho = n < nlo ? (nlo - x) / nlo : (x - (N - 1 - nhi)) / nhi;
                rho = rho > 1.0 ? 1.0 : (rho < 0.0 ? 0.0 : rho);
                double delta = npml * disc;
                double smax = -(c.order + 1) * std::log(c.R0) / (2.0 * delta);
                double rm = std::pow(rho, c.order);
                double sigma = smax * rm;
                double kappa = 1.0 + (c.kappa_max - 1.0) * rm;
                double alpha = c.alpha_max * (1.0 - rho);
                b = std::exp(-(sigma / kappa + alpha) * dt);
                double den = sigma * kappa + kappa * kappa * alpha;
                a = den > 0.0 ? sigma * (b - 1.0) / den : 0.0;
                kinv = 1.0 / kappa;
            }
            out[(3 * half + 0) * N + n] = (float)b;
            out[(3 * half + 1) * N + n] = (float)a;
            out[(3 * half + 2) * N + n] = (float)kinv;
        }
    }
}

}  // namespace

hipEvent_t timing_begin(fdtd_dev *d, int which) {
    KernelTiming &kt = d->ktime;
    if (kt.every <= 0 || (kt.launches[which]++ % kt.every) != 0) return nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    if (hipEventCreate(&ev_start) != hipSuccess || hipEventCreate(&ev_stop) != hipSuccess) return nullptr;
    if (hipEventRecord(ev_start, d->stream) != hipSuccess) return nullptr;
    kt.ev[which].push_back(ev_start);
    kt.ev[which].push_back(ev_stop);
    return ev_stop;
}

int cpml_tables(fdtd_dev *d, double dt) {
    Cpml &c = d->cpml;
    if (c.tab_dt == dt) return 0;
    const int N[3] = {d->I, d->J, d->K};
    std::vector<float> host((size_t)6 * (N[0] + N[1] + N[2]));
    for (int a = 0; a < 3; a++) cpml_profile(N[a], c.nlo[a], c.nhi[a], c, d->disc, dt, host.data() + c.tab_off[a]);
    FDTD_TRY(hipMemcpyAsync(c.tab, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice, d->stream));
    FDTD_TRY(hipStreamSynchronize(d->stream));
    c.tab_dt = dt;
    return 0;
}

namespace {

PmlArgs pml_args(fdtd_dev *d) {
    PmlArgs pm{};
    const Cpml &c = d->cpml;
    for (int a = 0; a < 3; a++) {
        pm.tab[a] = c.tab ? c.tab + c.tab_off[a] : nullptr;
        pm.nlo[a] = c.nlo[a];
        pm.nhi[a] = c.nhi[a];
        pm.W[a] = c.W[a];
        for (int q = 0; q < 2; q++) {
            pm.psiE[a][q] = c.psi[0][a][q];
            pm.psiH[a][q] = c.psi[1][a][q];
        }
    }
    return pm;
}

int launch_update(fdtd_dev *d, double dt, int k0, int k1, int scheme, bool e_field, bool pml) {
    int rc = ensure_coefficients(d, dt);
    if (rc) return rc;
    if (pml && (rc = cpml_tables(d, dt))) return rc;
    if (scheme != 0 && scheme != 1) {
        set_error("scheme must be 0 (semi-implicit) or 1 (explicit)");
        return -2;
    }
    k0 = k0 < 1 ? 1 : k0;
    k1 = k1 > d->K - 1 ? d->K - 1 : k1;
    if (k1 <= k0) return 0;
    YeeArgs a;
    a.Ex = d->F[EX];
    a.Ey = d->F[EY];
    a.Ez = d->F[EZ];
    a.Hx = d->F[HX];
    a.Hy = d->F[HY];
    a.Hz = d->F[HZ];
    a.mat = d->mat;
    a.coef = d->d_coef + (size_t)scheme * d->nmat * kCoefStride;
    a.I = d->I;
    a.J = d->J;
    a.K = d->K;
    a.P = d->P;
    a.k0 = k0;
    a.k1 = k1;
    a.plane_blocks = (int)((d->P + kBlock - 1) / kBlock);
    // enough z chunks that the launch holds several waves of blocks per CU
    int nz = k1 - k0;
    int target_chunks = (4096 + a.plane_blocks - 1) / a.plane_blocks;
    if (target_chunks < 1) target_chunks = 1;
    a.kchunk = (nz + target_chunks - 1) / target_chunks;
    if (a.kchunk < 4) a.kchunk = nz < 4 ? nz : 4;
    int zchunks = (nz + a.kchunk - 1) / a.kchunk;
    a.nsrc = 0;
    for (auto &sp : d->src) {
        a.src[a.nsrc].k = sp.k;
        a.src[a.nsrc].J = sp.J;
        a.nsrc++;
    }
    a.dense = d->dense_src;
    a.N = d->N;
    a.nrs = d->n_run_src;
    for (int q = 0; q < d->n_run_src; q++) a.rs[q] = d->run_src[q];
    PmlArgs pm = pml_args(d);
    dim3 grid((unsigned)(a.plane_blocks * zchunks));
    const bool u8 = d->mat_bytes == 1;
    hipEvent_t ev_stop = timing_begin(d, e_field ? 0 : 1);
#define FDTD_LAUNCH(KER)                                                                                  \
    do {                                                                                                  \
        if (u8 && pml) hipLaunchKernelGGL((KER<uint8_t, true>), grid, dim3(kBlock), 0, d->stream, a, pm);   \
        else if (u8) hipLaunchKernelGGL((KER<uint8_t, false>), grid, dim3(kBlock), 0, d->stream, a, pm);   \
        else if (pml) hipLaunchKernelGGL((KER<uint16_t, true>), grid, dim3(kBlock), 0, d->stream, a, pm);  \
        else hipLaunchKernelGGL((KER<uint16_t, false>), grid, dim3(kBlock), 0, d->stream, a, pm);          \
    } while (0)
    if (e_field) FDTD_LAUNCH(k_update_E);
    else FDTD_LAUNCH(k_update_H);
#undef FDTD_LAUNCH
    FDTD_TRY(hipGetLastError());
    if (ev_stop) FDTD_TRY(hipEventRecord(ev_stop, d->stream));
    return 0;
}

int launch_mur(fdtd_dev *d, double dt, int k0, int k1, int field) {
    int rc = ensure_coefficients(d, dt);
    if (rc) return rc;
    d->alt_valid = false;  // wall cells change: the fused step's second buffer must be resynced
    k0 = k0 < 0 ? 0 : k0;
    k1 = k1 > d->K ? d->K : k1;
    FaceArgs fa = face_args(d, 0);
    const int N[3] = {d->I, d->J, d->K};
    // face order x0, y0, z0, x_end, y_end, z_end: later faces overwrite edges and corners
    for (int side = 0; side < 2; side++) {
        for (int a = 0; a < 3; a++) {
            if (a == 2 && side == 0 && k0 != 0) continue;
            if (a == 2 && side == 1 && k1 != d->K) continue;
            int b = (a + 1) % 3, cc = (a + 2) % 3;
            long long nu = (b == 2) ? (k1 - k0) : N[b];
            long long nv = (cc == 2) ? (k1 - k0) : N[cc];
            long long n = nu * nv;
            if (n <= 0) continue;
            unsigned nb = (unsigned)((n + kBlock - 1) / kBlock);
            if (d->mat_bytes == 1)
                hipLaunchKernelGGL(k_mur_face<uint8_t>, dim3(nb), dim3(kBlock), 0, d->stream, fa, a, side, field, k0, k1);
            else
                hipLaunchKernelGGL(k_mur_face<uint16_t>, dim3(nb), dim3(kBlock), 0, d->stream, fa, a, side, field, k0,
                                   k1);
            FDTD_TRY(hipGetLastError());
        }
    }
    return 0;
}

}  // namespace

int cpml_zero(fdtd_dev *d) {
    Cpml &c = d->cpml;
    long long n = c.psi_n[0] + c.psi_n[1] + c.psi_n[2];
    if (c.psi_base && n > 0) FDTD_TRY(hipMemsetAsync(c.psi_base, 0, (size_t)4 * n * sizeof(float), d->stream));
    if (c.psi_alt && n > 0) FDTD_TRY(hipMemsetAsync(c.psi_alt, 0, (size_t)4 * n * sizeof(float), d->stream));
    return 0;
}

}  // namespace fdtd

using namespace fdtd;

extern "C" {

int fdtd_dev_snapshot_shells(fdtd_dev *d) {
    FaceArgs fa = face_args(d, 0);
    const long long a0 = (long long)d->J * d->K, a1 = (long long)d->K * d->I, a2 = (long long)d->I * d->J;
    long long total = 4 * (a0 + a1 + a2);
    unsigned nb = (unsigned)((total + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_snapshot, dim3(nb), dim3(kBlock), 0, d->stream, fa, total, a0, a1, a2);
    FDTD_TRY(hipGetLastError());
    return 0;
}

int fdtd_dev_set_shells(fdtd_dev *d, const double *old) {
    // host AoS old_fields -> shell buffer (only the tangential pairs of each face's two layers)
    const int N[3] = {d->I, d->J, d->K};
    long long total = d->shell_off[5] + 8LL * d->I * d->J;
    std::vector<float> sh((size_t)total);
    for (int f = 0; f < 6; f++) {
        int a = f % 3, side = f / 3, b = (a + 1) % 3, cc = (a + 2) % 3;
        long long area = (long long)N[b] * N[cc];
        for (int layer = 0; layer < 2; layer++)
            for (int v = 0; v < N[cc]; v++)
                for (int u = 0; u < N[b]; u++) {
                    int x[3];
                    x[a] = side == 0 ? layer : N[a] - 1 - layer;
                    x[b] = u;
                    x[cc] = v;
                    long long c = x[0] + (long long)N[0] * (x[1] + (long long)N[1] * x[2]);
                    float *s = &sh[d->shell_off[f] + layer * 4 * area + u + (long long)N[b] * v];
                    s[0] = (float)old[6 * c + b];
                    s[area] = (float)old[6 * c + cc];
                    s[2 * area] = (float)old[6 * c + 3 + b];
                    s[3 * area] = (float)old[6 * c + 3 + cc];
                }
    }
    FDTD_TRY(hipMemcpyAsync(d->shell, sh.data(), sh.size() * sizeof(float), hipMemcpyHostToDevice, d->stream));
    FDTD_TRY(hipStreamSynchronize(d->stream));
    return 0;
}

int fdtd_dev_update_E(fdtd_dev *d, double dt, int k0, int k1, int scheme) {
    return launch_update(d, dt, k0, k1, scheme, true, false);
}

int fdtd_dev_update_H(fdtd_dev *d, double dt, int k0, int k1, int scheme) {
    return launch_update(d, dt, k0, k1, scheme, false, false);
}

int fdtd_dev_update_E_cpml(fdtd_dev *d, double dt) {
    if (!d->cpml.on) {
        set_error("CPML is not enabled on this grid");
        return -2;
    }
    return launch_update(d, dt, 1, d->K - 1, 0, true, true);
}

int fdtd_dev_update_H_cpml(fdtd_dev *d, double dt) {
    if (!d->cpml.on) {
        set_error("CPML is not enabled on this grid");
        return -2;
    }
    return launch_update(d, dt, 1, d->K - 1, 0, false, true);
}

int fdtd_dev_mur_E(fdtd_dev *d, double dt, int k0, int k1) { return launch_mur(d, dt, k0, k1, 0); }

int fdtd_dev_mur_H(fdtd_dev *d, double dt, int k0, int k1) { return launch_mur(d, dt, k0, k1, 1); }

int fdtd_dev_step(fdtd_dev *d, double dt, int scheme) {
    int rc;
    if (d->cpml.on && scheme == 0) {
        if (d->fused_rows > 0 && fused_supported(d)) return fused_step(d, dt, d->fused_rows, d->fused_kchunk);
        if ((rc = fdtd_dev_update_E_cpml(d, dt))) return rc;
        return fdtd_dev_update_H_cpml(d, dt);
    }
    if ((rc = fdtd_dev_snapshot_shells(d))) return rc;
    if ((rc = fdtd_dev_update_E(d, dt, 1, d->K - 1, scheme))) return rc;
    if ((rc = fdtd_dev_mur_E(d, dt, 0, d->K))) return rc;
    if ((rc = fdtd_dev_update_H(d, dt, 1, d->K - 1, scheme))) return rc;
    if ((rc = fdtd_dev_mur_H(d, dt, 0, d->K))) return rc;
    return 0;
}

int fdtd_dev_cpml_on(fdtd_dev *d, const int npml[6], int order, double R0, double kappa_max, double alpha_max) {
    const int N[3] = {d->I, d->J, d->K};
    for (int a = 0; a < 3; a++) {
        if (npml[2 * a] < 0 || npml[2 * a + 1] < 0 || npml[2 * a] + npml[2 * a + 1] + 2 > N[a]) {
            set_error("PML thickness does not fit the grid");
            return -2;
        }
    }
    if (!(R0 > 0 && R0 < 1) || order < 0 || kappa_max < 1 || alpha_max < 0) {
        set_error("invalid PML grading parameters");
        return -2;
    }
    FDTD_TRY(hipStreamSynchronize(d->stream));
    fdtd_dev_cpml_off(d);
    Cpml &c = d->cpml;
    c.order = order;
    c.R0 = R0;
    c.kappa_max = kappa_max;
    c.alpha_max = alpha_max;
    long long off = 0;
    for (int a = 0; a < 3; a++) {
        c.nlo[a] = npml[2 * a];
        c.nhi[a] = npml[2 * a + 1];
        c.W[a] = c.nlo[a] + (c.nhi[a] > 0 ? c.nhi[a] + 1 : 0);
        c.tab_off[a] = off;
        off += 6LL * N[a];
        c.psi_n[a] = (long long)c.W[a] * N[(a + 1) % 3] * N[(a + 2) % 3];
    }
    FDTD_TRY(hipMalloc(&c.tab, (size_t)off * sizeof(float)));
    long long total = c.psi_n[0] + c.psi_n[1] + c.psi_n[2];
    if (total > 0) {
        FDTD_TRY(hipMalloc(&c.psi_base, (size_t)4 * total * sizeof(float)));
        FDTD_TRY(hipMemsetAsync(c.psi_base, 0, (size_t)4 * total * sizeof(float), d->stream));
    }
    float *p = c.psi_base;
    for (int a = 0; a < 3; a++)
        for (int f = 0; f < 2; f++)
            for (int q = 0; q < 2; q++) {
                c.psi[f][a][q] = c.psi_n[a] > 0 ? p : nullptr;
                p += c.psi_n[a];
            }
    c.tab_dt = -1.0;
    c.on = true;
    FDTD_TRY(hipStreamSynchronize(d->stream));
    return 0;
}

int fdtd_dev_kernel_timing(fdtd_dev *d, int every) {
    int cnt[3];
    double ms[3];
    int rc = fdtd_dev_kernel_times(d, cnt, ms);  // drop any pending sample
    if (rc) return rc;
    d->ktime.every = every < 0 ? 0 : every;
    for (int w = 0; w < 3; w++) d->ktime.launches[w] = 0;
    return 0;
}

int fdtd_dev_kernel_times(fdtd_dev *d, int counts[3], double ms_out[3]) {
    KernelTiming &kt = d->ktime;
    FDTD_TRY(hipStreamSynchronize(d->stream));
    for (int w = 0; w < 3; w++) {
        double sum = 0;
        int n = 0;
        for (size_t q = 0; q + 1 < kt.ev[w].size(); q += 2) {
            float ms = 0;
            FDTD_TRY(hipEventElapsedTime(&ms, kt.ev[w][q], kt.ev[w][q + 1]));
            sum += ms;
            n++;
        }
        for (hipEvent_t e : kt.ev[w]) (void)hipEventDestroy(e);
        kt.ev[w].clear();
        counts[w] = n;
        ms_out[w] = sum;
    }
    return 0;
}

int fdtd_dev_cpml_off(fdtd_dev *d) {
    Cpml &c = d->cpml;
    if (c.tab || c.psi_base) FDTD_TRY(hipStreamSynchronize(d->stream));
    (void)hipFree(c.tab);
    (void)hipFree(c.psi_base);
    (void)hipFree(c.psi_alt);
    c = Cpml();
    return 0;
}

}  // extern "C"
