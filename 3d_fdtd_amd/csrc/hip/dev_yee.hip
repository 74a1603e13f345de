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

#include "dev_onepass.h"
#include "dev_stencil.h"

namespace fdtd {
namespace {

// NT: the own-cell field loads carry the streaming hint (grids beyond the last-level cache).
template <typename MatT, bool PML, int KC, typename Off, bool NT>
__global__ __launch_bounds__(kBlock) void k_update_E(YeeArgs a, PmlArgs pm) {
    update_E_block<MatT, PML, KC, Off, NT>(a, pm, xcd_logical_block(blockIdx.x, gridDim.x));
}

template <typename MatT, bool PML, int KC, typename Off, bool NT>
__global__ __launch_bounds__(kBlock) void k_update_H(YeeArgs a, PmlArgs pm) {
    update_H_block<MatT, PML, false, KC, Off, NT>(a, pm, xcd_logical_block(blockIdx.x, gridDim.x));
}

// The monitor arguments raise the scalar register count of the runtime-chunk variant just past
// the 8-waves-per-SIMD budget; the attribute keeps the occupancy of the plain update there (a few
// scalars spill to vector lanes).  The unrolled variants are limited by vector registers anyway.
template <typename MatT, bool PML, int KC, typename Off, bool NT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(KC == 0 ? 8 : 1))) void k_update_H_mon(
    YeeArgs a, PmlArgs pm) {
    update_H_block<MatT, PML, true, KC, Off, NT>(a, pm, xcd_logical_block(blockIdx.x, gridDim.x));
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

// Enumeration axes of face a: ua runs fastest in the shell and across neighbouring threads, so it
// is the face's tangential axis that is fastest in memory (x on the y faces, where the cyclic
// order would put z first and make every thread touch another plane); va is the other one.  The
// shell keeps the two tangential components in the cyclic order b = (a + 1) % 3, cc = (a + 2) % 3
// whatever the enumeration.
__host__ __device__ __forceinline__ void face_enum_axes(int a, int *ua, int *va) {
    *ua = a == 1 ? 0 : (a + 1) % 3;
    *va = a == 1 ? 2 : (a + 2) % 3;
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
    int ua, va;
    face_enum_axes(a, &ua, &va);
    // the two layers of one face cell are neighbouring threads (neighbouring words on x faces)
    const int layer = (int)(r & 1);
    const long long uv = r >> 1;
    int x[3];
    x[a] = side == 0 ? layer : fa.N[a] - 1 - layer;
    x[ua] = (int)(uv % fa.N[ua]);
    x[va] = (int)(uv / fa.N[ua]);
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
    int ua, va;
    face_enum_axes(a, &ua, &va);
    const long long area = (long long)fa.N[b] * fa.N[cc];
    // enumerate only the z-restricted part of the face
    int ulo = 0, uhi = fa.N[ua], vlo = 0, vhi = fa.N[va];
    if (ua == 2) { ulo = k0; uhi = k1; }
    if (va == 2) { vlo = k0; vhi = k1; }
    const long long nu = uhi - ulo;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nu * (vhi - vlo)) return;
    const int u = ulo + (int)(t % nu), v = vlo + (int)(t / nu);
    int xb[3], xi[3];
    xb[a] = side == 0 ? 0 : fa.N[a] - 1;
    xi[a] = side == 0 ? 1 : fa.N[a] - 2;
    xb[ua] = xi[ua] = u;
    xb[va] = xi[va] = v;
    const long long cb = cell_of(xb, fa.N), ci = cell_of(xi, fa.N);
    const long long uv = u + (long long)fa.N[ua] * v;
    const int f = a + 3 * side;
    const float *sh0 = fa.shell + fa.shell_off[f] + uv;          // boundary layer
    const float *sh1 = sh0 + 4 * area;                          // inner layer
    const float *cf = fa.coef + (long long)mat_at<MatT>(fa.mat, cb) * kCoefStride + 18 + 6 * field + 2 * a;
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

// All six Mur faces of one field in one launch, with the result of the reference's sequential
// face order (x0, y0, z0, x_end, y_end, z_end; later faces overwrite the edge and corner values of
// earlier ones, and read the values earlier faces wrote next to them).
//   One thread per (face, face cell); of the cell's two tangential components it writes those whose
// last writer in that order is its face.  That value is the face's formula applied to the inner
// neighbour's value as it stood when the face ran: if an earlier face covers the inner neighbour
// with the component tangential, that face's formula at the neighbour (and so on inward: a cell
// lies on at most three faces, so the chain has at most three levels), else the neighbour's value
// in memory.  Such a base value is never written by any face (were a face to write it, the same
// face would also cover this chain's first cell later in the order), so the threads do not race,
// and the float operations are those of the per-face kernels in the same order.
// Element i of (v0, v1, v2) by selects: indexing a small local array with a runtime index would
// place it in scratch memory.
template <typename T>
__device__ __forceinline__ T sel3(int i, T v0, T v1, T v2) {
    return i == 0 ? v0 : (i == 1 ? v1 : v2);
}

template <typename T>
__device__ __forceinline__ T sel6(int i, const T *v) {
    return i < 3 ? sel3(i, v[0], v[1], v[2]) : sel3(i - 3, v[3], v[4], v[5]);
}

// Face g (0..5 = x0, y0, z0, x_end, y_end, z_end) covers cell (x0, x1, x2) of an N0 x N1 x N2 grid.
__device__ __forceinline__ bool on_face(int x0, int x1, int x2, int g, int N0, int N1, int N2) {
    const int ag = g % 3;
    return sel3(ag, x0, x1, x2) == (g < 3 ? 0 : sel3(ag, N0, N1, N2) - 1);
}

// The latest face before `before` that covers (x0, x1, x2) with component m tangential, or -1.
__device__ __forceinline__ int earlier_face(int x0, int x1, int x2, int before, int m, int N0, int N1, int N2) {
    for (int g = before - 1; g >= 0; g--)
        if (g % 3 != m && on_face(x0, x1, x2, g, N0, N1, N2)) return g;
    return -1;
}

// Cell (x0, x1, x2) moved one cell inward along face g's normal.
__device__ __forceinline__ void inward(int g, int N0, int N1, int N2, int &x0, int &x1, int &x2) {
    const int ag = g % 3, v = g < 3 ? 1 : sel3(ag, N0, N1, N2) - 2;
    x0 = ag == 0 ? v : x0;
    x1 = ag == 1 ? v : x1;
    x2 = ag == 2 ? v : x2;
}

// Face g's Mur update of component m at its boundary cell (x0, x1, x2), given the inner
// neighbour's current value R: sh1 + s (R - sh0) from the shells and the cell's coefficient.
template <typename MatT>
__device__ __forceinline__ float mur_formula(const FaceArgs &fa, long long a0, long long a1, long long a2, int field,
                                             int g, int m, int x0, int x1, int x2, float R) {
    const int N0 = fa.N[0], N1 = fa.N[1];
    const int ag = g % 3, qg = m == (ag + 1) % 3 ? 0 : 1;
    // tangential enumeration axes (face_enum_axes): x faces (z... y, z), y faces (x, z), z faces (x, y)
    const int u = ag == 0 ? x1 : x0, v = ag == 2 ? x1 : x2, Nu = ag == 0 ? N1 : N0;
    const long long ar = sel3(ag, a0, a1, a2);
    const float *sh0 = fa.shell + sel6(g, fa.shell_off) + u + (long long)Nu * v;
    const float *sh1 = sh0 + 4 * ar;
    const long long cl = (long long)x0 + (long long)N0 * ((long long)x1 + (long long)N1 * x2);
    const float cf = fa.coef[(long long)mat_at<MatT>(fa.mat, cl) * kCoefStride + 18 + 6 * field + 2 * ag + qg];
    return sh1[(2 * field + qg) * ar] + cf * (R - sh0[(2 * field + qg) * ar]);
}

// All six Mur faces of one field in one launch, with the result of the reference's sequential
// face order (x0, y0, z0, x_end, y_end, z_end; later faces overwrite the edge and corner values of
// earlier ones, and read the values earlier faces wrote next to them).
//   One thread per (face, face cell); of the cell's two tangential components it writes those whose
// last writer in that order is its face.  That value is the face's formula applied to the inner
// neighbour's value as it stood when the face ran: if an earlier face covers the inner neighbour
// with the component tangential, that face's formula at the neighbour (and so on inward: each level
// moves one more coordinate inward, so the chain has at most three levels), else the neighbour's
// value in memory.  Such a base value is never written by any face (were a face to write it, the
// same face would also cover this chain's first cell later in the order), so the threads do not
// race, and the float operations are those of the per-face kernels in the same order.
template <typename MatT>
__global__ void k_mur_all(FaceArgs fa, int field, long long a0, long long a1, long long a2) {
    const int N0 = fa.N[0], N1 = fa.N[1], N2 = fa.N[2];
    long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int f = 0;
    while (f < 6 && r >= sel3(f % 3, a0, a1, a2)) {
        r -= sel3(f % 3, a0, a1, a2);
        f++;
    }
    if (f >= 6) return;
    const int a = f % 3;
    // enumeration (face_enum_axes): u fastest in memory -- y on x faces, x on y and z faces
    const int Nu = a == 0 ? N1 : N0;
    const int u = (int)(r % Nu), v = (int)(r / Nu);
    const int bnd = f < 3 ? 0 : sel3(a, N0, N1, N2) - 1;
    const int b0 = a == 0 ? bnd : u;
    const int b1 = a == 1 ? bnd : (a == 0 ? u : v);
    const int b2 = a == 2 ? bnd : v;
    const long long cb = (long long)b0 + (long long)N0 * ((long long)b1 + (long long)N1 * b2);
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const int m = (a + 1 + q) % 3;
        bool last = true;
        for (int g = f + 1; g < 6; g++)
            if (g % 3 != m && on_face(b0, b1, b2, g, N0, N1, N2)) last = false;
        if (!last) continue;
        // level 1 (the inner neighbour of this cell) and level 2 (its inner neighbour)
        int c0 = b0, c1 = b1, c2 = b2;
        inward(f, N0, N1, N2, c0, c1, c2);
        const int g1 = earlier_face(c0, c1, c2, f, m, N0, N1, N2);
        int d0 = c0, d1 = c1, d2 = c2, g2 = -1;
        if (g1 >= 0) {
            inward(g1, N0, N1, N2, d0, d1, d2);
            g2 = earlier_face(d0, d1, d2, g1, m, N0, N1, N2);
        }
        int e0 = d0, e1 = d1, e2 = d2;
        if (g2 >= 0) inward(g2, N0, N1, N2, e0, e1, e2);
        // base: the neighbour beyond the last level, then the formulas from the innermost out
        const int z0 = g1 < 0 ? c0 : (g2 < 0 ? d0 : e0);
        const int z1 = g1 < 0 ? c1 : (g2 < 0 ? d1 : e1);
        const int z2 = g1 < 0 ? c2 : (g2 < 0 ? d2 : e2);
        float *Fm = sel6(3 * field + m, fa.F);
        float R = Fm[(long long)z0 + (long long)N0 * ((long long)z1 + (long long)N1 * z2)];
        if (g2 >= 0) R = mur_formula<MatT>(fa, a0, a1, a2, field, g2, m, d0, d1, d2, R);
        if (g1 >= 0) R = mur_formula<MatT>(fa, a0, a1, a2, field, g1, m, c0, c1, c2, R);
        R = mur_formula<MatT>(fa, a0, a1, a2, field, f, m, b0, b1, b2, R);
        Fm[cb] = R;
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

void timing_events(fdtd_dev *d, int which, hipEvent_t *start, hipEvent_t *stop) {
    KernelTiming &kt = d->ktime;
    *start = *stop = nullptr;
    if (kt.every <= 0 || (kt.launches[which]++ % kt.every) != 0) return;
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess) return;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return;
    }
    kt.ev[which].push_back(a);
    kt.ev[which].push_back(b);
    *start = a;
    *stop = b;
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

// Off: cell-offset type of the kernels (uint32_t for grids below 2^30 cells, else long long).
// nt: streaming loads of the own-cell fields (PML grids beyond the last-level cache only).
// mb: material id bytes; 4-byte ids (per-cell media) take the runtime-chunk kernels only (KC 0),
// without the streaming hint.
template <int KC, typename Off>
void launch_split(bool e_field, bool mon, int mb, bool ext, bool pml, bool nt, dim3 grid, hipStream_t st,
                  hipEvent_t t0, hipEvent_t t1, const YeeArgs &a, const PmlArgs &pm) {
#define FDTD_GO(KER) hipExtLaunchKernelGGL(KER, grid, dim3(kBlock), 0, st, t0, t1, 0, a, pm)
#define FDTD_PML(KERN, MT)                                  \
    do {                                                    \
        if (nt) FDTD_GO((KERN<MT, true, KC, Off, true>));   \
        else FDTD_GO((KERN<MT, true, KC, Off, false>));     \
    } while (0)
    if constexpr (KC == 0) {
        if (mb == 4) {
            if (e_field && pml) FDTD_GO((k_update_E<uint32_t, true, 0, Off, false>));
            else if (e_field) FDTD_GO((k_update_E<uint32_t, false, 0, Off, false>));
            else if (!mon && pml) FDTD_GO((k_update_H<uint32_t, true, 0, Off, false>));
            else if (!mon) FDTD_GO((k_update_H<uint32_t, false, 0, Off, false>));
            else if (pml) FDTD_GO((k_update_H_mon<uint32_t, true, 0, Off, false>));
            else FDTD_GO((k_update_H_mon<uint32_t, false, 0, Off, false>));
            return;
        }
    }
    if (ext && pml) {  // extruded ids from the K x I table (PML grids: the batched and run paths)
        if (e_field) FDTD_PML(k_update_E, ExtIds);
        else if (!mon) FDTD_PML(k_update_H, ExtIds);
        else FDTD_PML(k_update_H_mon, ExtIds);
        return;
    }
    const bool u8 = mb == 1;
    if (e_field) {
        if (u8 && pml) FDTD_PML(k_update_E, uint8_t);
        else if (u8) FDTD_GO((k_update_E<uint8_t, false, KC, Off, false>));
        else if (pml) FDTD_PML(k_update_E, uint16_t);
        else FDTD_GO((k_update_E<uint16_t, false, KC, Off, false>));
    } else if (!mon) {
        if (u8 && pml) FDTD_PML(k_update_H, uint8_t);
        else if (u8) FDTD_GO((k_update_H<uint8_t, false, KC, Off, false>));
        else if (pml) FDTD_PML(k_update_H, uint16_t);
        else FDTD_GO((k_update_H<uint16_t, false, KC, Off, false>));
    } else {
        if (u8 && pml) FDTD_PML(k_update_H_mon, uint8_t);
        else if (u8) FDTD_GO((k_update_H_mon<uint8_t, false, KC, Off, false>));
        else if (pml) FDTD_PML(k_update_H_mon, uint16_t);
        else FDTD_GO((k_update_H_mon<uint16_t, false, KC, Off, false>));
    }
#undef FDTD_PML
#undef FDTD_GO
}

// Launch arguments of the split kernels over [k0, k1) (no run sources or monitors).
int base_args(fdtd_dev *d, double dt, int k0, int k1, int scheme, bool pml, YeeArgs *ap, PmlArgs *pmp) {
    int rc = ensure_coefficients(d, dt);
    if (rc) return rc;
    if (pml && (rc = cpml_tables(d, dt))) return rc;
    if (scheme != 0 && scheme != 1) {
        set_error("scheme must be 0 (semi-implicit) or 1 (explicit)");
        return -2;
    }
    YeeArgs &a = *ap;
    a.Ex = d->F[EX];
    a.Ey = d->F[EY];
    a.Ez = d->F[EZ];
    a.Hx = d->F[HX];
    a.Hy = d->F[HY];
    a.Hz = d->F[HZ];
    a.mat = d->mat;
    a.coef = d->d_coef + (size_t)scheme * d->nmat * kCoefStride;
    a.ext = d->ext_ids;
    a.ext_j0 = d->ext_j0;
    a.ext_j1 = d->ext_j1;
    a.ext_bg = d->ext_bg;
    a.I = d->I;
    a.J = d->J;
    a.K = d->K;
    a.P = d->P;
    a.k0 = k0;
    a.k1 = k1;
    a.plane_blocks = (int)((d->P + kBlock - 1) / kBlock);
    // The kernels are latency-bound: two z planes per thread, both loaded before either is
    // computed, with many resident wavefronts measured fastest on MI355X (strip waveguide
    // 120x100x240: 2 planes 51.6, 3 planes 48.5, 5 planes 47.8 G cell-updates/s, same run).
    int nz = k1 - k0;
    a.kchunk = nz < 2 ? nz : 2;
    if (d->split_kchunk > 0) a.kchunk = d->split_kchunk < nz ? d->split_kchunk : nz;
    a.nsrc = 0;
    for (auto &sp : d->src) {
        a.src[a.nsrc].k = sp.k;
        a.src[a.nsrc].J = sp.J;
        a.nsrc++;
    }
    a.dense = d->dense_src;
    a.N = d->N;
    a.nrs = 0;
    a.nmon = 0;
    a.step = nullptr;
    a.step0 = 0;
    set_source_range(a);
    *pmp = pml_args(d);
    return 0;
}

int launch_update(fdtd_dev *d, double dt, int k0, int k1, int scheme, bool e_field, bool pml) {
    fields_changed(d);
    k0 = k0 < 1 ? 1 : k0;
    k1 = k1 > d->K - 1 ? d->K - 1 : k1;
    if (k1 <= k0) return ensure_coefficients(d, dt);
    YeeArgs a;
    PmlArgs pm;
    int rc = base_args(d, dt, k0, k1, scheme, pml, &a, &pm);
    if (rc) return rc;
    const int nz = k1 - k0;
    const int zchunks = (nz + a.kchunk - 1) / a.kchunk;
    a.nrs = d->n_run_src;
    for (int q = 0; q < d->n_run_src; q++) a.rs[q] = d->run_src[q];
    set_source_range(a);
    // run monitors ride on a full-range H update only
    const bool mon = !e_field && d->n_run_mon > 0 && k0 == 1 && k1 == d->K - 1;
    a.nmon = mon ? d->n_run_mon : 0;
    for (int q = 0; q < a.nmon; q++) a.mon[q] = d->run_mon[q];
    if (mon) d->n_run_mon = 0;  // consumed
    dim3 grid((unsigned)(a.plane_blocks * zchunks));
    const int mb = d->mat_bytes;
    const bool nt = pml && streaming_loads(d->N, d->stream_mode);
    const bool ext = d->ext_ids != nullptr;
    hipEvent_t t0, t1;
    timing_events(d, e_field ? 0 : 1, &t0, &t1);
    if (d->N < kSmallGridCells && mb != 4) {
        switch (a.kchunk) {
            case 2: launch_split<2, uint32_t>(e_field, mon, mb, ext, pml, nt, grid, d->stream, t0, t1, a, pm); break;
            case 3: launch_split<3, uint32_t>(e_field, mon, mb, ext, pml, nt, grid, d->stream, t0, t1, a, pm); break;
            case 4: launch_split<4, uint32_t>(e_field, mon, mb, ext, pml, nt, grid, d->stream, t0, t1, a, pm); break;
            default: launch_split<0, uint32_t>(e_field, mon, mb, ext, pml, nt, grid, d->stream, t0, t1, a, pm); break;
        }
    } else if (d->N < kSmallGridCells) {
        launch_split<0, uint32_t>(e_field, mon, mb, false, pml, nt, grid, d->stream, t0, t1, a, pm);
    } else {
        launch_split<0, long long>(e_field, mon, mb, false, pml, nt, grid, d->stream, t0, t1, a, pm);
    }
    FDTD_TRY(hipGetLastError());
    return 0;
}

int launch_mur(fdtd_dev *d, double dt, int k0, int k1, int field) {
    fields_changed(d);
    int rc = ensure_coefficients(d, dt);
    if (rc) return rc;
    d->alt_valid = false;  // wall cells change: the fused step's second buffer must be resynced
    k0 = k0 < 0 ? 0 : k0;
    k1 = k1 > d->K ? d->K : k1;
    FaceArgs fa = face_args(d, 0);
    const int N[3] = {d->I, d->J, d->K};
    if (k0 == 0 && k1 == d->K && d->I >= 3 && d->J >= 3 && d->K >= 3) {
        // the whole boundary: one launch for the six faces
        const long long ar[3] = {(long long)N[1] * N[2], (long long)N[2] * N[0], (long long)N[0] * N[1]};
        const long long total = 2 * (ar[0] + ar[1] + ar[2]);
        const unsigned nb = (unsigned)((total + kBlock - 1) / kBlock);
        if (d->mat_bytes == 1)
            hipLaunchKernelGGL(k_mur_all<uint8_t>, dim3(nb), dim3(kBlock), 0, d->stream, fa, field, ar[0], ar[1], ar[2]);
        else if (d->mat_bytes == 2)
            hipLaunchKernelGGL(k_mur_all<uint16_t>, dim3(nb), dim3(kBlock), 0, d->stream, fa, field, ar[0], ar[1],
                               ar[2]);
        else
            hipLaunchKernelGGL(k_mur_all<uint32_t>, dim3(nb), dim3(kBlock), 0, d->stream, fa, field, ar[0], ar[1],
                               ar[2]);
        FDTD_TRY(hipGetLastError());
        return 0;
    }
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
            else if (d->mat_bytes == 2)
                hipLaunchKernelGGL(k_mur_face<uint16_t>, dim3(nb), dim3(kBlock), 0, d->stream, fa, a, side, field, k0,
                                   k1);
            else
                hipLaunchKernelGGL(k_mur_face<uint32_t>, dim3(nb), dim3(kBlock), 0, d->stream, fa, a, side, field, k0,
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

bool step_uses_split(const fdtd_dev *d) {
    if (!d->cpml.on || d->mat_bytes == 4) return true;
    if (d->fused_rows > 0) return !fused_supported(d);
    if (d->fused_rows == -1) return !col_supported(d);
    if (d->fused_rows == -3) return !onepass_supported(d);
    return true;  // split pair, or leap launches (a lone step is a split pair; runs use leap_run)
}

bool step_uses_onepass(const fdtd_dev *d) {
    return d->cpml.on && d->mat_bytes <= 2 && d->fused_rows == -3 && onepass_supported(d);
}

bool step_folds_monitors(const fdtd_dev *d) {
    return step_uses_split(d) || step_uses_onepass(d) || (d->cpml.on && d->fused_rows == -1 && col_supported(d));
}

int step_monitor_slots(const fdtd_dev *d) {
    if (step_uses_onepass(d))
        return onepass_plane_slots(d, d->onepass_tw > 0 ? d->onepass_tw : onepass_default_width(d));
    return (int)((d->P + kBlock - 1) / kBlock);
}

int fill_step_args(fdtd_dev *d, double dt, YeeArgs *a, PmlArgs *pm) {
    if (d->K < 3) {
        set_error("grid too thin for a step");
        return -2;
    }
    return base_args(d, dt, 1, d->K - 1, 0, d->cpml.on, a, pm);
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
        int ua, va;
        face_enum_axes(a, &ua, &va);
        long long area = (long long)N[b] * N[cc];
        for (int layer = 0; layer < 2; layer++)
            for (int v = 0; v < N[va]; v++)
                for (int u = 0; u < N[ua]; u++) {
                    int x[3];
                    x[a] = side == 0 ? layer : N[a] - 1 - layer;
                    x[ua] = u;
                    x[va] = v;
                    long long c = x[0] + (long long)N[0] * (x[1] + (long long)N[1] * x[2]);
                    float *s = &sh[d->shell_off[f] + layer * 4 * area + u + (long long)N[ua] * v];
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
        if (d->fused_rows == -1 && col_supported(d)) return col_step(d, dt, d->fused_kchunk);
        if (step_uses_onepass(d))
            return onepass_step(d, dt, d->onepass_tw > 0 ? d->onepass_tw : onepass_default_width(d),
                                d->onepass_kc > 0 ? d->onepass_kc : onepass_default_chunk(d));
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
    FDTD_TRY(dev_malloc(&c.tab, (size_t)off * sizeof(float)));
    long long total = c.psi_n[0] + c.psi_n[1] + c.psi_n[2];
    if (total > 0) {
        FDTD_TRY(dev_malloc(&c.psi_base, (size_t)4 * total * sizeof(float)));
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
    (void)dev_free(c.tab);
    (void)dev_free(c.psi_base);
    (void)dev_free(c.psi_alt);
    c = Cpml();
    return 0;
}

}  // extern "C"
