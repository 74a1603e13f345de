// Device grid lifecycle, materials, host<->device transfers, timers.
// Replaces the storage side of timefield3d.c:6-136 with device-resident FP32 structure-of-arrays
// fields and a material-id volume (see fdtd3d_dev.h for the C-ABI contract).
#include <cmath>
#include <cstring>
#include <mutex>
#include <unordered_map>

#include "dev_internal.h"

namespace fdtd {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

int fail(const char *where, hipError_t e) {
    set_error(std::string(where) + ": " + hipGetErrorString(e));
    return -1;
}

namespace {

struct MatKey {
    float v[12];
    bool operator==(const MatKey &o) const { return std::memcmp(v, o.v, sizeof v) == 0; }
};

struct MatKeyHash {
    size_t operator()(const MatKey &k) const {
        uint64_t h = 1469598103934665603ull;
        const unsigned char *p = reinterpret_cast<const unsigned char *>(k.v);
        for (size_t i = 0; i < sizeof k.v; i++) h = (h ^ p[i]) * 1099511628211ull;
        return (size_t)h;
    }
};

// Material ids of an extruded geometry: rows j0 <= j < j1 take ids[k][i], every other cell the
// background id (the structure of the reference's waveguide and taper setters,
// timefield3d.c:1458-1535, and of adjoint3d_taper_insert).
template <typename T>
__global__ void k_extrude_ids(T *__restrict__ mat, const uint16_t *__restrict__ ids, int I, int J, long long N,
                              int j0, int j1, int background) {
    long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const long long P = (long long)I * J;
    const int k = (int)(c / P);
    const long long p = c - (long long)k * P;
    const int j = (int)(p / I), i = (int)(p - (long long)j * I);
    mat[c] = (T)(j >= j0 && j < j1 ? ids[(long long)k * I + i] : background);
}

// Region of a repaint: x in [lo, hi) of rows y in [y0, y1), z in [z0, z1).  For a linear taper the
// x range of plane k is [x0 + round(ls (k - z0)), x1 + round(rs (k - z0))), evaluated in double as
// the reference evaluates it (timefield3d.c:1511-1535); for a prism it is [x0, x1).  Cells outside
// the grid are skipped.
struct PaintArgs {
    int taper;
    int x0, x1, y0, y1, z0, z1;  // z0, z1: planes of the launch (inside the grid)
    int zs;                      // taper: plane where the slopes start (start_z)
    double ls, rs;
};

// Cells of the region take material remap[id] (the same tuple with the new permittivity and
// permeability); one block per (row, 256 columns).
template <typename T>
__global__ void k_paint(T *__restrict__ mat, const uint32_t *__restrict__ remap, int I, int J, PaintArgs p,
                        int xblocks) {
    const long long row = blockIdx.x / xblocks;
    const int i = (int)(blockIdx.x % xblocks) * blockDim.x + threadIdx.x;
    const int ny = p.y1 - p.y0;
    const int j = p.y0 + (int)(row % ny), k = p.z0 + (int)(row / ny);
    int lo = p.x0;
    double hi = p.x1;
    if (p.taper) {
        lo = (int)(p.x0 + round(p.ls * (k - p.zs)));
        hi = p.x1 + round(p.rs * (k - p.zs));
    }
    if (i < lo || i < 0 || i >= I || !((double)i < hi)) return;
    const long long c = i + (long long)I * (j + (long long)J * k);
    mat[c] = (T)remap[mat[c]];
}

// Wider material ids (the table outgrew 1 or 2 bytes).
template <typename S, typename T>
__global__ void k_widen_ids(const S *__restrict__ src, T *__restrict__ dst, long long N) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < N) dst[c] = (T)src[c];
}

// AoS double[6N] -> six SoA float arrays.
__global__ void k_aos_to_soa(const double *__restrict__ aos, float *__restrict__ base, long long N, long long Nal) {
    long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
#pragma unroll
    for (int m = 0; m < 6; m++) base[m * Nal + c] = (float)aos[6 * c + m];
}

__global__ void k_soa_to_aos(const float *__restrict__ base, double *__restrict__ aos, long long N, long long Nal) {
    long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
#pragma unroll
    for (int m = 0; m < 6; m++) aos[6 * c + m] = (double)base[m * Nal + c];
}

// The id volume is about to change: the extruded table no longer describes it (no kernel still
// queued reads it: batched calls end synchronised, and the single-grid kernels read the volume).
void drop_extruded_ids(fdtd_dev *d) {
    (void)dev_free(d->ext_ids);
    d->ext_ids = nullptr;
}

}  // namespace

int ensure_coefficients(fdtd_dev *d, double dt) {
    if (!d->have_mats) {
        set_error("materials were never uploaded");
        return -1;
    }
    if (d->coef_dt == dt) return 0;
    std::vector<float> tab(coef_table_floats(d->nmat), 0.0f);
    const double inv_dx = 1 / d->disc;
    for (int id = 0; id < d->nmat; id++) {
        const float *pr = &d->props[(size_t)id * kPropStride];
        for (int scheme = 0; scheme < 2; scheme++) {
            float *r = &tab[((size_t)scheme * d->nmat + id) * kCoefStride];
            for (int m = 0; m < 3; m++) {
                double ep = pr[m], mu = pr[3 + m], se = pr[6 + m], sm = pr[9 + m];
                if (scheme == 0) {
                    // semi-implicit (timefield3d.c:1823-1833 and 1743-1753; sigma* counted twice)
                    double inv_e = 1.0 / (ep / dt + 0.5 * se), keep_e = ep / dt - 0.5 * se;
                    double inv_h = 1.0 / (mu / dt + 0.5 * sm), keep_h = mu / dt - 0.5 * sm;
                    r[0 + m] = (float)(inv_e * keep_e);
                    r[3 + m] = (float)(inv_e * inv_dx);
                    r[6 + m] = (float)inv_e;
                    r[9 + m] = (float)(inv_h * (keep_h - sm));
                    r[12 + m] = (float)(inv_h * inv_dx);
                    r[15 + m] = (float)inv_h;
                } else {
                    // explicit (timefield3d.c:1593-1603 and 1664-1674)
                    r[0 + m] = (float)(1.0 - dt * se / ep);
                    r[3 + m] = (float)(dt * inv_dx / ep);
                    r[6 + m] = (float)(dt / ep);
                    r[9 + m] = (float)(1.0 - dt * sm / mu);
                    r[12 + m] = (float)(dt * inv_dx / mu);
                    r[15 + m] = (float)(dt / mu);
                }
            }
            // Mur: E comp t on a face normal to axis a pairs ep_t with mu_d, H comp t pairs
            // ep_d with mu_t, d the third axis (timefield3d.c:2275-2721).
            for (int a = 0; a < 3; a++) {
                for (int q = 0; q < 2; q++) {
                    int t = (a + 1 + q) % 3, dd = 3 - a - t;
                    double vE = 1 / std::sqrt((double)pr[t] * (double)pr[3 + dd]);
                    double vH = 1 / std::sqrt((double)pr[dd] * (double)pr[3 + t]);
                    r[18 + 2 * a + q] = (float)((vE * dt - d->disc) / (vE * dt + d->disc));
                    r[24 + 2 * a + q] = (float)((vH * dt - d->disc) / (vH * dt + d->disc));
                }
            }
        }
        const float *s = &tab[(size_t)id * kCoefStride];  // semi-implicit record
        float *pk = &tab[coef_packed_offset(d->nmat) + (size_t)id * kCoefPacked];
        for (int m = 0; m < 3; m++) {
            pk[m] = s[m];           // caE
            pk[3 + m] = s[3 + m];   // cbE
            pk[6 + m] = s[9 + m];   // caH
            pk[9 + m] = s[12 + m];  // cbH
        }
        tab[coef_iso_offset(d->nmat) + id] = s[3];
    }
    // isotropic lossless table: caE = caH = 1, cbE equal over the components, cbH one value
    bool iso = d->nmat > 0;
    const float cbh = d->nmat > 0 ? tab[12] : 0.0f;
    for (int id = 0; id < d->nmat && iso; id++) {
        const float *s = &tab[(size_t)id * kCoefStride];
        for (int m = 0; m < 3; m++)
            iso = iso && s[m] == 1.0f && s[9 + m] == 1.0f && s[3 + m] == s[3] && s[12 + m] == cbh;
    }
    d->coef_iso = iso;
    d->coef_cbh = cbh;
    FDTD_TRY(hipMemcpyAsync(d->d_coef, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice, d->stream));
    FDTD_TRY(hipStreamSynchronize(d->stream));
    d->coef_dt = dt;
    return 0;
}

const double *mode_on_device(fdtd_dev *d, const double *mode, int mi, int mj, double mdisc, int *err) {
    *err = 0;
    const size_t n = (size_t)12 * mi * mj;
    d->use_clock++;
    for (auto &mc : d->modes) {
        if (mc.host_key != mode || mc.mi != mi || mc.mj != mj || mc.mdisc != mdisc) continue;
        // Re-upload whenever the caller's mode changed (in-place renormalisation, mode_smul).
        mc.last_use = d->use_clock;
        if (std::memcmp(mode, mc.shadow.data(), n * sizeof(double)) == 0) return mc.dev;
        std::memcpy(mc.shadow.data(), mode, n * sizeof(double));
        mc.version = ++d->mode_uploads;
        hipError_t e = hipMemcpyAsync(mc.dev, mode, n * sizeof(double), hipMemcpyHostToDevice, d->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
        if (e != hipSuccess) {
            *err = fail("mode upload", e);
            return nullptr;
        }
        return mc.dev;
    }
    // new entry; keep at most 16 cached modes per grid (evict the least recently used)
    if (d->modes.size() >= 16) {
        size_t victim = 0;
        for (size_t q = 1; q < d->modes.size(); q++)
            if (d->modes[q].last_use < d->modes[victim].last_use) victim = q;
        (void)dev_free(d->modes[victim].dev);
        d->modes.erase(d->modes.begin() + victim);
    }
    ModeCache mc;
    mc.host_key = mode;
    mc.mi = mi;
    mc.mj = mj;
    mc.mdisc = mdisc;
    mc.shadow.assign(mode, mode + n);
    mc.last_use = d->use_clock;
    mc.version = ++d->mode_uploads;
    hipError_t e = dev_malloc(&mc.dev, n * sizeof(double));
    if (e == hipSuccess) e = hipMemcpyAsync(mc.dev, mode, n * sizeof(double), hipMemcpyHostToDevice, d->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    if (e != hipSuccess) {
        *err = fail("mode upload", e);
        return nullptr;
    }
    d->modes.push_back(std::move(mc));
    return d->modes.back().dev;
}

const ModeCache *mode_cached(const fdtd_dev *d, const double *mode, int mi, int mj, double mdisc) {
    for (const auto &mc : d->modes)
        if (mc.host_key == mode && mc.mi == mi && mc.mj == mj && mc.mdisc == mdisc) return &mc;
    return nullptr;
}

// Rewrite the id volume with wider ids (the table outgrew them, or a batch needs one width).
int widen_material_ids(fdtd_dev *d, int bytes) {
    if (bytes <= d->mat_bytes) return 0;
    FDTD_TRY(hipStreamSynchronize(d->stream));  // kernels already enqueued may still read the ids
    const unsigned nbN = (unsigned)((d->N + kBlock - 1) / kBlock);
    void *w = nullptr;
    FDTD_TRY(dev_malloc(&w, (size_t)d->N * bytes));
    if (d->mat_bytes == 1 && bytes == 2)
        hipLaunchKernelGGL((k_widen_ids<uint8_t, uint16_t>), dim3(nbN), dim3(kBlock), 0, d->stream,
                           static_cast<const uint8_t *>(d->mat), static_cast<uint16_t *>(w), d->N);
    else if (d->mat_bytes == 1)
        hipLaunchKernelGGL((k_widen_ids<uint8_t, uint32_t>), dim3(nbN), dim3(kBlock), 0, d->stream,
                           static_cast<const uint8_t *>(d->mat), static_cast<uint32_t *>(w), d->N);
    else
        hipLaunchKernelGGL((k_widen_ids<uint16_t, uint32_t>), dim3(nbN), dim3(kBlock), 0, d->stream,
                           static_cast<const uint16_t *>(d->mat), static_cast<uint32_t *>(w), d->N);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    if (e != hipSuccess) {
        (void)dev_free(w);
        return fail("widen material ids", e);
    }
    (void)dev_free(d->mat);
    d->mat = w;
    d->mat_bytes = bytes;
    return 0;
}

}  // namespace fdtd

using namespace fdtd;

extern "C" {

const char *fdtd_dev_last_error(void) { return g_err.c_str(); }

fdtd_dev *fdtd_dev_create(int I, int J, int K, double disc) {
    if (I < 3 || J < 3 || K < 3) {
        set_error("grid dimensions must be at least 3 in every axis");
        return nullptr;
    }
    fdtd_dev *d = new fdtd_dev();
    d->I = I;
    d->J = J;
    d->K = K;
    d->P = (long long)I * J;
    d->N = d->P * K;
    d->disc = disc;
    d->nmat = 0;
    d->mat_bytes = 1;
    d->mat = nullptr;
    d->d_props = nullptr;
    d->d_coef = nullptr;
    d->coef_dt = NAN;
    d->coef_iso = false;
    d->coef_cbh = 0.0f;
    d->have_mats = false;
    d->fourier = nullptr;
    d->partials = nullptr;
    d->red_out = nullptr;
    d->ovl_count = nullptr;
    d->ovl_host = nullptr;
    d->ovl_dev = nullptr;
    d->field_epoch = 1;
    d->mode_uploads = 0;
    d->n_seen = d->n_prev = d->n_batch = 0;
    d->seen_epoch = d->batch_epoch = 0;
    d->partial_cap = 0;
    d->use_clock = 0;
    d->field_base = nullptr;
    d->shell = nullptr;
    d->dense_src = nullptr;
    d->run = nullptr;
    d->alt_base = nullptr;
    d->alt_valid = false;
    d->n_run_src = 0;
    d->n_run_mon = 0;
    d->fused_rows = 0;  // split E/H kernels: measured faster than the fused step on MI355X
    d->fused_kchunk = 8;
    d->split_kchunk = 0;
    d->leap_kc = 0;
    d->onepass_tw = 0;
    d->onepass_kc = 0;
    d->finite_every = 0;
    d->finite_buf = nullptr;
    d->stream_mode = -1;
    d->ev0 = d->ev1 = d->xev = nullptr;
    hipError_t e = dev_stream_acquire(&d->stream);
    // fields: 6 arrays, each N floats, each starting on a 256-byte boundary
    d->Nal = (d->N + 63) / 64 * 64;
    if (e == hipSuccess) e = dev_malloc(&d->field_base, (size_t)6 * d->Nal * sizeof(float));
    if (e == hipSuccess) e = hipMemsetAsync(d->field_base, 0, (size_t)6 * d->Nal * sizeof(float), d->stream);
    for (int m = 0; m < 6; m++) d->F[m] = d->field_base + m * d->Nal;
    // shells: faces x0,y0,z0,xend,yend,zend; each 2 layers x 4 components x area
    const long long area[3] = {(long long)J * K, (long long)K * I, (long long)I * J};
    long long off = 0;
    for (int f = 0; f < 6; f++) {
        d->shell_off[f] = off;
        off += 8 * area[f % 3];
    }
    if (e == hipSuccess) e = dev_malloc(&d->shell, (size_t)off * sizeof(float));
    if (e == hipSuccess) e = hipMemsetAsync(d->shell, 0, (size_t)off * sizeof(float), d->stream);
    if (e == hipSuccess) e = dev_malloc(&d->red_out, sizeof(double2) * 4);
    if (e == hipSuccess) e = dev_malloc(&d->ovl_count, kOvlSlots * sizeof(unsigned));
    if (e == hipSuccess) e = hipMemsetAsync(d->ovl_count, 0, kOvlSlots * sizeof(unsigned), d->stream);
    if (e == hipSuccess) e = hipHostMalloc((void **)&d->ovl_host, kOvlSlots * sizeof(double2), hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void **)&d->ovl_dev, d->ovl_host, 0);
    if (e == hipSuccess) e = hipEventCreate(&d->ev0);
    if (e == hipSuccess) e = hipEventCreate(&d->ev1);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&d->xev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    if (e != hipSuccess) {
        fail("fdtd_dev_create", e);
        fdtd_dev_destroy(d);
        return nullptr;
    }
    return d;
}

void fdtd_dev_destroy(fdtd_dev *d) {
    if (!d) return;
    if (d->stream) (void)hipStreamSynchronize(d->stream);
    (void)dev_free(d->field_base);
    (void)dev_free(d->alt_base);
    (void)dev_free(d->mat);
    (void)dev_free(d->ext_ids);
    (void)dev_free(d->d_props);
    (void)dev_free(d->d_coef);
    (void)dev_free(d->shell);
    (void)dev_free(d->dense_src);
    fdtd_dev_cpml_off(d);
    run_release(d);
    (void)dev_free(d->fourier);
    (void)dev_free(d->finite_buf);
    (void)dev_free(d->partials);
    (void)dev_free(d->red_out);
    (void)dev_free(d->ovl_count);
    if (d->ovl_host) (void)hipHostFree(d->ovl_host);
    for (auto &p : d->src) (void)dev_free(p.J);
    for (auto &m : d->modes) (void)dev_free(m.dev);
    if (d->ev0) (void)hipEventDestroy(d->ev0);
    if (d->xev) (void)hipEventDestroy(d->xev);
    if (d->ev1) (void)hipEventDestroy(d->ev1);
    if (d->stream) (void)dev_stream_release(d->stream);
    delete d;
}

int fdtd_dev_set_materials(fdtd_dev *d, const float *ep, const float *mu, const float *sig) {
    // kernels already enqueued may still read the id volume and the tables
    FDTD_TRY(hipStreamSynchronize(d->stream));
    std::unordered_map<MatKey, int, MatKeyHash> ids;
    std::vector<uint32_t> cell_id((size_t)d->N);
    std::vector<float> props;
    static const float kOne3[3] = {1.0f, 1.0f, 1.0f};
    static const float kZero6[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    MatKey prev;
    int prev_id = -1;
    for (long long c = 0; c < d->N; c++) {
        // runs of equal cells are the common case: compare in place before building a key
        const float *e3 = ep + 3 * c, *m3 = mu ? mu + 3 * c : kOne3, *s6 = sig ? sig + 6 * c : kZero6;
        if (prev_id >= 0 && std::memcmp(e3, prev.v, 3 * sizeof(float)) == 0 &&
            std::memcmp(m3, prev.v + 3, 3 * sizeof(float)) == 0 && std::memcmp(s6, prev.v + 6, 6 * sizeof(float)) == 0) {
            cell_id[c] = (uint32_t)prev_id;
            continue;
        }
        MatKey k;
        std::memcpy(k.v, e3, 3 * sizeof(float));
        std::memcpy(k.v + 3, m3, 3 * sizeof(float));
        std::memcpy(k.v + 6, s6, 6 * sizeof(float));
        int id;
        auto it = ids.find(k);
        if (it == ids.end()) {
            id = (int)ids.size();
            ids.emplace(k, id);
            props.resize((size_t)(id + 1) * kPropStride, 0.0f);
            std::memcpy(&props[(size_t)id * kPropStride], k.v, sizeof k.v);
        } else {
            id = it->second;
        }
        prev = k;
        prev_id = id;
        cell_id[c] = (uint32_t)id;
    }
    // ids of 1, 2 or 4 bytes: a per-cell graded medium (every cell its own tuple) takes 4-byte ids
    // and a table as long as the grid, and runs on the split kernels
    int nmat = (int)ids.size();
    int bytes = nmat <= 256 ? 1 : (nmat <= 65536 ? 2 : 4);
    drop_extruded_ids(d);
    if (d->mat && bytes != d->mat_bytes) {
        FDTD_TRY(dev_free(d->mat));
        d->mat = nullptr;
    }
    if (!d->mat) FDTD_TRY(dev_malloc(&d->mat, (size_t)d->N * bytes));
    if (bytes == 1) {
        std::vector<uint8_t> c8((size_t)d->N);
        for (long long c = 0; c < d->N; c++) c8[c] = (uint8_t)cell_id[c];
        FDTD_TRY(hipMemcpy(d->mat, c8.data(), (size_t)d->N, hipMemcpyHostToDevice));
    } else if (bytes == 2) {
        std::vector<uint16_t> c16((size_t)d->N);
        for (long long c = 0; c < d->N; c++) c16[c] = (uint16_t)cell_id[c];
        FDTD_TRY(hipMemcpy(d->mat, c16.data(), (size_t)d->N * 2, hipMemcpyHostToDevice));
    } else {
        FDTD_TRY(hipMemcpy(d->mat, cell_id.data(), (size_t)d->N * 4, hipMemcpyHostToDevice));
    }
    if (nmat != d->nmat || !d->d_props) {
        (void)dev_free(d->d_props);
        (void)dev_free(d->d_coef);
        d->d_props = nullptr;
        d->d_coef = nullptr;
        FDTD_TRY(dev_malloc(&d->d_props, (size_t)nmat * kPropStride * sizeof(float)));
        FDTD_TRY(dev_malloc(&d->d_coef, coef_table_floats(nmat) * sizeof(float)));
    }
    FDTD_TRY(hipMemcpy(d->d_props, props.data(), props.size() * sizeof(float), hipMemcpyHostToDevice));
    d->props = std::move(props);
    d->nmat = nmat;
    d->mat_bytes = bytes;
    d->coef_dt = NAN;
    d->have_mats = true;
    d->mat_uploads++;
    return 0;
}

namespace {

// Material ids -> per-cell AoS ep[3], mu[3], sig[6] (any of them NULL to skip).
int expand_materials(fdtd_dev *d, float *ep, float *mu, float *sig) {
    FDTD_TRY(hipStreamSynchronize(d->stream));
    std::vector<unsigned char> raw((size_t)d->N * d->mat_bytes);
    FDTD_TRY(hipMemcpy(raw.data(), d->mat, raw.size(), hipMemcpyDeviceToHost));
    for (long long c = 0; c < d->N; c++) {
        const size_t id = d->mat_bytes == 1   ? raw[c]
                          : d->mat_bytes == 2 ? reinterpret_cast<const uint16_t *>(raw.data())[c]
                                              : reinterpret_cast<const uint32_t *>(raw.data())[c];
        const float *pr = &d->props[id * kPropStride];
        for (int m = 0; m < 3; m++) {
            if (ep) ep[3 * c + m] = pr[m];
            if (mu) mu[3 * c + m] = pr[3 + m];
        }
        if (sig)
            for (int m = 0; m < 6; m++) sig[6 * c + m] = pr[6 + m];
    }
    return 0;
}

// Repaint a region on the device: each cell keeps its conductivities and takes permittivity ep
// (and permeability mu when set_mu) on all three components.  The host maps every table entry to
// its repainted tuple (appending the new ones), widens the id volume when the table outgrows it,
// and a kernel rewrites the ids of the region's cells: no per-cell upload.
int paint(fdtd_dev *d, PaintArgs p, float ep, float mu, bool set_mu) {
    int rc;
    if (!d->have_mats) {
        set_error("paint: the grid has no materials yet");
        return -2;
    }
    if (p.y0 < 0) p.y0 = 0;
    if (p.y1 > d->J) p.y1 = d->J;
    if (p.z0 < 0) p.z0 = 0;
    if (p.z1 > d->K) p.z1 = d->K;
    if (p.y1 <= p.y0 || p.z1 <= p.z0) return 0;
    FDTD_TRY(hipStreamSynchronize(d->stream));  // kernels already enqueued may still read the ids
    std::unordered_map<MatKey, int, MatKeyHash> ids;
    for (int m = 0; m < d->nmat; m++) {
        MatKey k;
        std::memcpy(k.v, &d->props[(size_t)m * kPropStride], sizeof k.v);
        ids.emplace(k, m);
    }
    const int nmat0 = d->nmat;
    std::vector<uint32_t> remap((size_t)nmat0);
    std::vector<float> props = d->props;
    for (int m = 0; m < nmat0; m++) {
        MatKey k;
        std::memcpy(k.v, &d->props[(size_t)m * kPropStride], sizeof k.v);
        for (int q = 0; q < 3; q++) {
            k.v[q] = ep;
            if (set_mu) k.v[3 + q] = mu;
        }
        auto it = ids.find(k);
        if (it == ids.end()) {
            const int id = (int)ids.size();
            it = ids.emplace(k, id).first;
            props.resize((size_t)(id + 1) * kPropStride, 0.0f);
            std::memcpy(&props[(size_t)id * kPropStride], k.v, sizeof k.v);
        }
        remap[m] = (uint32_t)it->second;
    }
    const int nmat = (int)ids.size();
    const int bytes = nmat <= 256 ? 1 : (nmat <= 65536 ? 2 : 4);
    if (bytes > d->mat_bytes && (rc = widen_material_ids(d, bytes))) return rc;
    drop_extruded_ids(d);
    uint32_t *drm = nullptr;
    FDTD_TRY(dev_malloc(&drm, remap.size() * sizeof(uint32_t)));
    hipError_t e = hipMemcpyAsync(drm, remap.data(), remap.size() * sizeof(uint32_t), hipMemcpyHostToDevice, d->stream);
    if (e == hipSuccess) {
        const int xb = (d->I + kBlock - 1) / kBlock;
        const unsigned nb = (unsigned)((long long)(p.y1 - p.y0) * (p.z1 - p.z0) * xb);
        if (d->mat_bytes == 1)
            hipLaunchKernelGGL(k_paint<uint8_t>, dim3(nb), dim3(kBlock), 0, d->stream, static_cast<uint8_t *>(d->mat),
                               drm, d->I, d->J, p, xb);
        else if (d->mat_bytes == 2)
            hipLaunchKernelGGL(k_paint<uint16_t>, dim3(nb), dim3(kBlock), 0, d->stream,
                               static_cast<uint16_t *>(d->mat), drm, d->I, d->J, p, xb);
        else
            hipLaunchKernelGGL(k_paint<uint32_t>, dim3(nb), dim3(kBlock), 0, d->stream,
                               static_cast<uint32_t *>(d->mat), drm, d->I, d->J, p, xb);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    (void)dev_free(drm);
    if (e != hipSuccess) return fail("paint", e);
    if (nmat != nmat0) {
        (void)dev_free(d->d_props);
        (void)dev_free(d->d_coef);
        d->d_props = nullptr;
        d->d_coef = nullptr;
        FDTD_TRY(dev_malloc(&d->d_props, (size_t)nmat * kPropStride * sizeof(float)));
        FDTD_TRY(dev_malloc(&d->d_coef, coef_table_floats(nmat) * sizeof(float)));
        FDTD_TRY(hipMemcpy(d->d_props, props.data(), props.size() * sizeof(float), hipMemcpyHostToDevice));
        d->props = std::move(props);
        d->nmat = nmat;
        d->coef_dt = NAN;
    }
    d->mat_paints++;
    return 0;
}

}  // namespace

int fdtd_dev_paint_taper(fdtd_dev *d, float ep, float mu, int start_left, int start_right, int end_left,
                         int end_right, int start_z, int end_z, int bottom, int top) {
    PaintArgs p{};
    p.taper = 1;
    p.x0 = start_left;
    p.x1 = start_right;
    p.y0 = bottom;
    p.y1 = top;
    p.z0 = start_z;
    p.z1 = end_z;
    p.zs = start_z;
    if (end_z > start_z) {
        p.rs = (double)(end_right - start_right) / ((double)(end_z - start_z));
        p.ls = (double)(end_left - start_left) / ((double)(end_z - start_z));
    }
    return paint(d, p, ep, mu, true);
}

int fdtd_dev_paint_prism(fdtd_dev *d, float ep, float mu, int set_mu, int x0, int x1, int y0, int y1, int z0, int z1) {
    PaintArgs p{};
    p.x0 = x0 < 0 ? 0 : x0;
    p.x1 = x1 > d->I ? d->I : x1;
    p.y0 = y0;
    p.y1 = y1;
    p.z0 = z0;
    p.z1 = z1;
    if (p.x1 <= p.x0) return 0;
    return paint(d, p, ep, mu, set_mu != 0);
}

int fdtd_dev_get_materials(fdtd_dev *d, float *ep, float *mu, float *sig) {
    if (!d->have_mats) {
        set_error("get_materials: the grid has no materials yet");
        return -2;
    }
    return expand_materials(d, ep, mu, sig);
}

int fdtd_dev_material_counts(fdtd_dev *d, long long counts[3]) {
    counts[0] = d->mat_uploads;
    counts[1] = d->mat_paints;
    counts[2] = d->nmat;
    return 0;
}

int fdtd_dev_set_materials_extruded(fdtd_dev *d, int nmat, const float *ep, const float *mu, const float *sig,
                                    const uint16_t *ids, int j0, int j1, int background) {
    if (nmat < 1 || nmat > 65536 || !ep || !ids || background < 0 || background >= nmat || j0 < 0 || j1 > d->J ||
        j0 > j1) {
        set_error("set_materials_extruded: bad table size, id, row range or null array");
        return -2;
    }
    const long long nid = (long long)d->K * d->I;
    for (long long q = 0; q < nid; q++)
        if (ids[q] >= nmat) {
            set_error("set_materials_extruded: material id outside the table");
            return -2;
        }
    // kernels already enqueued may still read the id volume and the tables
    FDTD_TRY(hipStreamSynchronize(d->stream));
    std::vector<float> props((size_t)nmat * kPropStride, 0.0f);
    for (int m = 0; m < nmat; m++) {
        float *pr = &props[(size_t)m * kPropStride];
        for (int q = 0; q < 3; q++) {
            pr[q] = ep[3 * m + q];
            pr[3 + q] = mu ? mu[3 * m + q] : 1.0f;
        }
        for (int q = 0; q < 6; q++) pr[6 + q] = sig ? sig[6 * m + q] : 0.0f;
    }
    const int bytes = nmat <= 256 ? 1 : 2;
    if (d->mat && bytes != d->mat_bytes) {
        FDTD_TRY(dev_free(d->mat));
        d->mat = nullptr;
    }
    if (!d->mat) FDTD_TRY(dev_malloc(&d->mat, (size_t)d->N * bytes));
    uint16_t *dids = nullptr;
    FDTD_TRY(dev_malloc(&dids, (size_t)nid * sizeof(uint16_t)));
    hipError_t e = hipMemcpyAsync(dids, ids, (size_t)nid * sizeof(uint16_t), hipMemcpyHostToDevice, d->stream);
    if (e == hipSuccess) {
        const unsigned nb = (unsigned)((d->N + kBlock - 1) / kBlock);
        if (bytes == 1)
            hipLaunchKernelGGL(k_extrude_ids<uint8_t>, dim3(nb), dim3(kBlock), 0, d->stream,
                               static_cast<uint8_t *>(d->mat), dids, d->I, d->J, d->N, j0, j1, background);
        else
            hipLaunchKernelGGL(k_extrude_ids<uint16_t>, dim3(nb), dim3(kBlock), 0, d->stream,
                               static_cast<uint16_t *>(d->mat), dids, d->I, d->J, d->N, j0, j1, background);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);  // the host id array may go right after
    drop_extruded_ids(d);
    if (e != hipSuccess) {
        (void)dev_free(dids);
        return fail("fdtd_dev_set_materials_extruded", e);
    }
    d->ext_ids = dids;  // the volume now equals the table
    d->ext_j0 = j0;
    d->ext_j1 = j1;
    d->ext_bg = background;
    if (nmat != d->nmat || !d->d_props) {
        (void)dev_free(d->d_props);
        (void)dev_free(d->d_coef);
        d->d_props = nullptr;
        d->d_coef = nullptr;
        FDTD_TRY(dev_malloc(&d->d_props, (size_t)nmat * kPropStride * sizeof(float)));
        FDTD_TRY(dev_malloc(&d->d_coef, coef_table_floats(nmat) * sizeof(float)));
    }
    FDTD_TRY(hipMemcpy(d->d_props, props.data(), props.size() * sizeof(float), hipMemcpyHostToDevice));
    d->props = std::move(props);
    d->nmat = nmat;
    d->mat_bytes = bytes;
    d->coef_dt = NAN;
    d->have_mats = true;
    d->mat_uploads++;
    return 0;
}

int fdtd_dev_set_fields(fdtd_dev *d, const double *fields) {
    fields_changed(d);
    d->alt_valid = false;
    double *tmp = nullptr;
    FDTD_TRY(dev_malloc(&tmp, (size_t)6 * d->N * sizeof(double)));
    hipError_t e = hipMemcpyAsync(tmp, fields, (size_t)6 * d->N * sizeof(double), hipMemcpyHostToDevice, d->stream);
    if (e == hipSuccess) {
        unsigned nb = (unsigned)((d->N + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(k_aos_to_soa, dim3(nb), dim3(kBlock), 0, d->stream, tmp, d->field_base, d->N, d->Nal);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    (void)dev_free(tmp);
    if (e != hipSuccess) return fail("fdtd_dev_set_fields", e);
    return 0;
}

int fdtd_dev_get_fields(fdtd_dev *d, double *fields) {
    double *tmp = nullptr;
    FDTD_TRY(dev_malloc(&tmp, (size_t)6 * d->N * sizeof(double)));
    unsigned nb = (unsigned)((d->N + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_soa_to_aos, dim3(nb), dim3(kBlock), 0, d->stream, d->field_base, tmp, d->N, d->Nal);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(fields, tmp, (size_t)6 * d->N * sizeof(double), hipMemcpyDeviceToHost, d->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    (void)dev_free(tmp);
    if (e != hipSuccess) return fail("fdtd_dev_get_fields", e);
    return 0;
}

int fdtd_dev_set_sources(fdtd_dev *d, const float *sources) {
    // A z plane is kept on the device when any of its source values is non-zero.
    std::vector<int> planes;
    for (int k = 0; k < d->K; k++) {
        const float *p = sources + 6 * d->P * k;
        bool nz = false;
        for (long long q = 0; q < 6 * d->P && !nz; q++) nz = p[q] != 0.0f;
        if (nz) planes.push_back(k);
    }
    FDTD_TRY(hipStreamSynchronize(d->stream));
    for (auto &p : d->src) (void)dev_free(p.J);
    d->src.clear();
    (void)dev_free(d->dense_src);
    d->dense_src = nullptr;
    if ((int)planes.size() > kMaxSrcPlanes) {
        // sources spread over many planes: keep the whole 6 x N volume (extra stream per cell)
        std::vector<float> soa((size_t)6 * d->N);
        for (long long c = 0; c < d->N; c++)
            for (int m = 0; m < 6; m++) soa[m * d->N + c] = sources[6 * c + m];
        FDTD_TRY(dev_malloc(&d->dense_src, soa.size() * sizeof(float)));
        FDTD_TRY(hipMemcpy(d->dense_src, soa.data(), soa.size() * sizeof(float), hipMemcpyHostToDevice));
        return 0;
    }
    std::vector<float> soa((size_t)6 * d->P);
    for (int k : planes) {
        const float *p = sources + 6 * d->P * k;
        for (long long c = 0; c < d->P; c++)
            for (int m = 0; m < 6; m++) soa[m * d->P + c] = p[6 * c + m];
        SrcPlane sp;
        sp.k = k;
        sp.stride = d->P;
        FDTD_TRY(dev_malloc(&sp.J, (size_t)6 * d->P * sizeof(float)));
        FDTD_TRY(hipMemcpy(sp.J, soa.data(), soa.size() * sizeof(float), hipMemcpyHostToDevice));
        d->src.push_back(sp);
    }
    return 0;
}

int fdtd_dev_get_sources(fdtd_dev *d, float *sources) {
    std::memset(sources, 0, (size_t)6 * d->N * sizeof(float));
    FDTD_TRY(hipStreamSynchronize(d->stream));
    if (d->dense_src) {
        std::vector<float> soa((size_t)6 * d->N);
        FDTD_TRY(hipMemcpy(soa.data(), d->dense_src, soa.size() * sizeof(float), hipMemcpyDeviceToHost));
        for (long long c = 0; c < d->N; c++)
            for (int m = 0; m < 6; m++) sources[6 * c + m] = soa[m * d->N + c];
        return 0;
    }
    std::vector<float> soa((size_t)6 * d->P);
    for (auto &sp : d->src) {
        FDTD_TRY(hipMemcpy(soa.data(), sp.J, soa.size() * sizeof(float), hipMemcpyDeviceToHost));
        float *p = sources + 6 * d->P * sp.k;
        for (long long c = 0; c < d->P; c++)
            for (int m = 0; m < 6; m++) p[6 * c + m] = soa[m * d->P + c];
    }
    return 0;
}

int fdtd_dev_zero(fdtd_dev *d) {
    fields_changed(d);
    FDTD_TRY(hipMemsetAsync(d->field_base, 0, (size_t)6 * d->Nal * sizeof(float), d->stream));
    if (d->alt_base) {
        FDTD_TRY(hipMemsetAsync(d->alt_base, 0, (size_t)6 * d->Nal * sizeof(float), d->stream));
        d->alt_valid = true;
    }
    long long shell_n = d->shell_off[5] + 8LL * d->I * d->J;
    FDTD_TRY(hipMemsetAsync(d->shell, 0, (size_t)shell_n * sizeof(float), d->stream));
    for (auto &p : d->src) FDTD_TRY(hipMemsetAsync(p.J, 0, (size_t)6 * d->P * sizeof(float), d->stream));
    if (d->dense_src) FDTD_TRY(hipMemsetAsync(d->dense_src, 0, (size_t)6 * d->N * sizeof(float), d->stream));
    if (int rc = cpml_zero(d)) return rc;
    if (d->fourier) FDTD_TRY(hipMemsetAsync(d->fourier, 0, (size_t)6 * d->N * sizeof(double2), d->stream));
    return 0;
}

int fdtd_dev_get_fourier(fdtd_dev *d, double *out) {
    if (!d->fourier) {
        std::memset(out, 0, (size_t)12 * d->N * sizeof(double));
        return 0;
    }
    std::vector<double2> soa((size_t)6 * d->N);
    FDTD_TRY(hipStreamSynchronize(d->stream));
    FDTD_TRY(hipMemcpy(soa.data(), d->fourier, soa.size() * sizeof(double2), hipMemcpyDeviceToHost));
    for (long long c = 0; c < d->N; c++)
        for (int m = 0; m < 6; m++) {
            out[2 * (6 * c + m)] = soa[m * d->N + c].x;
            out[2 * (6 * c + m) + 1] = soa[m * d->N + c].y;
        }
    return 0;
}

int fdtd_dev_set_fourier(fdtd_dev *d, const double *in) {
    if (!d->fourier) FDTD_TRY(dev_malloc(&d->fourier, (size_t)6 * d->N * sizeof(double2)));
    std::vector<double2> soa((size_t)6 * d->N);
    for (long long c = 0; c < d->N; c++)
        for (int m = 0; m < 6; m++) soa[m * d->N + c] = make_double2(in[2 * (6 * c + m)], in[2 * (6 * c + m) + 1]);
    FDTD_TRY(hipStreamSynchronize(d->stream));
    FDTD_TRY(hipMemcpy(d->fourier, soa.data(), soa.size() * sizeof(double2), hipMemcpyHostToDevice));
    return 0;
}

int fdtd_dev_set_step_kernel(fdtd_dev *d, int rows, int kchunk) {
    if (rows == -2) {  // single-pass leap launches for device-resident runs
        if (kchunk < 0) {
            set_error("leap launches: z planes per chunk must be >= 0 (0 = automatic)");
            return -2;
        }
        d->fused_rows = -2;
        d->leap_kc = kchunk;
        return 0;
    }
    if (rows == -3) {  // single-pass full steps
        if (kchunk < 0) {
            set_error("single-pass step: z planes per chunk must be >= 0 (0 = automatic)");
            return -2;
        }
        d->fused_rows = -3;
        d->onepass_tw = 0;
        d->onepass_kc = kchunk;
        return 0;
    }
    if (rows != -1 && rows != 0 && rows != 1 && rows != 2 && rows != 4 && rows != 8 && rows != 16 && rows != 32) {
        set_error("step kernel rows must be -3 (single-pass full steps), -2 (single-pass leap launches), -1 (fused "
                  "columns), 0 (split kernels), 1, 2, 4, 8 (fused register strips) or 16, 32 (fused LDS tiles of 8 / 16 "
                  "rows)");
        return -2;
    }
    if (rows != 0 && kchunk < 1) {
        set_error("fused step z chunk must be positive");
        return -2;
    }
    if (rows == -1 && kchunk != 1 && kchunk != 2 && kchunk != 4) {
        set_error("fused column step: z planes per thread must be 1, 2 or 4");
        return -2;
    }
    d->fused_rows = rows;
    if (rows == 0) {
        d->split_kchunk = kchunk > 0 ? kchunk : 0;
    } else {
        d->fused_kchunk = kchunk;
    }
    return 0;
}

int fdtd_dev_set_streaming_loads(fdtd_dev *d, int mode) {
    if (mode < -1 || mode > 1) {
        set_error("streaming loads: mode must be -1 (automatic), 0 (off) or 1 (on)");
        return -2;
    }
    d->stream_mode = mode;
    return 0;
}

int fdtd_dev_timer_start(fdtd_dev *d) {
    FDTD_TRY(hipEventRecord(d->ev0, d->stream));
    return 0;
}

int fdtd_dev_timer_stop(fdtd_dev *d, double *elapsed_ms) {
    FDTD_TRY(hipEventRecord(d->ev1, d->stream));
    FDTD_TRY(hipEventSynchronize(d->ev1));
    float ms = 0;
    FDTD_TRY(hipEventElapsedTime(&ms, d->ev0, d->ev1));
    *elapsed_ms = ms;
    return 0;
}

int fdtd_dev_synchronize(fdtd_dev *d) {
    FDTD_TRY(hipStreamSynchronize(d->stream));
    return 0;
}

}  // extern "C"
