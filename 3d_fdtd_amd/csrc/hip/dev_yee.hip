// Yee leapfrog on gfx950: split E / H z-marching kernels, boundary-shell snapshot and the
// first-order Mur faces.  Reference semantics: timefield3d.c:246-346 (shell snapshot),
// 1537-1837 (E/H updates), 2275-2721 (Mur faces), 2723-2849 (step drivers).
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
};

template <typename MatT>
__device__ __forceinline__ int mat_id(const void *mat, long long c) {
    return (int)static_cast<const MatT *>(mat)[c];
}

__device__ __forceinline__ const float *src_plane(const YeeArgs &a, int k) {
    const float *J = nullptr;
    for (int s = 0; s < a.nsrc; s++)
        if (a.src[s].k == k) J = a.src[s].J;
    return J;
}

// E <- caE*E + cbE*(backward-difference curl of H) - cjE*J   (timefield3d.c:1759-1837)
template <typename MatT>
__global__ __launch_bounds__(kBlock) void k_update_E(YeeArgs a) {
    const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const int pb = lb % a.plane_blocks, kc = lb / a.plane_blocks;
    const long long p = (long long)pb * kBlock + threadIdx.x;
    if (p >= a.P) return;
    const int j = (int)(p / a.I), i = (int)(p - (long long)j * a.I);
    if (i < 1 || i > a.I - 2 || j < 1 || j > a.J - 2) return;
    const int kb = a.k0 + kc * a.kchunk;
    const int ke = min(a.k1, kb + a.kchunk);
    if (kb >= ke) return;
    long long c = p + (long long)kb * a.P;
    float hx_km = a.Hx[c - a.P], hy_km = a.Hy[c - a.P];
    for (int k = kb; k < ke; k++, c += a.P) {
        const float hx = a.Hx[c], hy = a.Hy[c], hz = a.Hz[c];
        const float hy_im = a.Hy[c - 1], hz_im = a.Hz[c - 1];
        const float hx_jm = a.Hx[c - a.I], hz_jm = a.Hz[c - a.I];
        const float *cf = a.coef + (long long)mat_id<MatT>(a.mat, c) * kCoefStride;
        const float dx = (hz - hz_jm) - (hy - hy_km);
        const float dy = (hx - hx_km) - (hz - hz_im);
        const float dz = (hy - hy_im) - (hx - hx_jm);
        float ex = cf[0] * a.Ex[c] + cf[3] * dx;
        float ey = cf[1] * a.Ey[c] + cf[4] * dy;
        float ez = cf[2] * a.Ez[c] + cf[5] * dz;
        const float *J = src_plane(a, k);
        if (J) {
            ex -= cf[6] * J[p];
            ey -= cf[7] * J[a.P + p];
            ez -= cf[8] * J[2 * a.P + p];
        }
        a.Ex[c] = ex;
        a.Ey[c] = ey;
        a.Ez[c] = ez;
        hx_km = hx;
        hy_km = hy;
    }
}

// H <- caH*H - cbH*(forward-difference curl of E) - cjH*M   (timefield3d.c:1679-1757)
template <typename MatT>
__global__ __launch_bounds__(kBlock) void k_update_H(YeeArgs a) {
    const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const int pb = lb % a.plane_blocks, kc = lb / a.plane_blocks;
    const long long p = (long long)pb * kBlock + threadIdx.x;
    if (p >= a.P) return;
    const int j = (int)(p / a.I), i = (int)(p - (long long)j * a.I);
    if (i < 1 || i > a.I - 2 || j < 1 || j > a.J - 2) return;
    const int kb = a.k0 + kc * a.kchunk;
    const int ke = min(a.k1, kb + a.kchunk);
    if (kb >= ke) return;
    long long c = p + (long long)kb * a.P;
    float ex = a.Ex[c], ey = a.Ey[c];
    for (int k = kb; k < ke; k++, c += a.P) {
        const float ez = a.Ez[c];
        const float ex_kp = a.Ex[c + a.P], ey_kp = a.Ey[c + a.P];
        const float ey_ip = a.Ey[c + 1], ez_ip = a.Ez[c + 1];
        const float ex_jp = a.Ex[c + a.I], ez_jp = a.Ez[c + a.I];
        const float *cf = a.coef + (long long)mat_id<MatT>(a.mat, c) * kCoefStride;
        const float dx = (ez_jp - ez) - (ey_kp - ey);
        const float dy = (ex_kp - ex) - (ez_ip - ez);
        const float dz = (ey_ip - ey) - (ex_jp - ex);
        float hx = cf[9] * a.Hx[c] - cf[12] * dx;
        float hy = cf[10] * a.Hy[c] - cf[13] * dy;
        float hz = cf[11] * a.Hz[c] - cf[14] * dz;
        const float *J = src_plane(a, k);
        if (J) {
            hx -= cf[15] * J[3 * a.P + p];
            hy -= cf[16] * J[4 * a.P + p];
            hz -= cf[17] * J[5 * a.P + p];
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

int launch_update(fdtd_dev *d, double dt, int k0, int k1, int scheme, bool e_field) {
    int rc = ensure_coefficients(d, dt);
    if (rc) return rc;
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
    dim3 grid((unsigned)(a.plane_blocks * zchunks));
    if (e_field) {
        if (d->mat_bytes == 1) hipLaunchKernelGGL(k_update_E<uint8_t>, grid, dim3(kBlock), 0, d->stream, a);
        else hipLaunchKernelGGL(k_update_E<uint16_t>, grid, dim3(kBlock), 0, d->stream, a);
    } else {
        if (d->mat_bytes == 1) hipLaunchKernelGGL(k_update_H<uint8_t>, grid, dim3(kBlock), 0, d->stream, a);
        else hipLaunchKernelGGL(k_update_H<uint16_t>, grid, dim3(kBlock), 0, d->stream, a);
    }
    FDTD_TRY(hipGetLastError());
    return 0;
}

int launch_mur(fdtd_dev *d, double dt, int k0, int k1, int field) {
    int rc = ensure_coefficients(d, dt);
    if (rc) return rc;
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
    return launch_update(d, dt, k0, k1, scheme, true);
}

int fdtd_dev_update_H(fdtd_dev *d, double dt, int k0, int k1, int scheme) {
    return launch_update(d, dt, k0, k1, scheme, false);
}

int fdtd_dev_mur_E(fdtd_dev *d, double dt, int k0, int k1) { return launch_mur(d, dt, k0, k1, 0); }

int fdtd_dev_mur_H(fdtd_dev *d, double dt, int k0, int k1) { return launch_mur(d, dt, k0, k1, 1); }

int fdtd_dev_step(fdtd_dev *d, double dt, int scheme) {
    int rc;
    if ((rc = fdtd_dev_snapshot_shells(d))) return rc;
    if ((rc = fdtd_dev_update_E(d, dt, 1, d->K - 1, scheme))) return rc;
    if ((rc = fdtd_dev_mur_E(d, dt, 0, d->K))) return rc;
    if ((rc = fdtd_dev_update_H(d, dt, 1, d->K - 1, scheme))) return rc;
    if ((rc = fdtd_dev_mur_H(d, dt, 0, d->K))) return rc;
    return 0;
}

}  // extern "C"
