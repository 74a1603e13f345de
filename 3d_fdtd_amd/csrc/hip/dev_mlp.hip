// Transmission surrogate: a small MLP (n_in -> width x depth with GELU -> n_out) trained and
// evaluated on the GPU with hand-written BF16 MFMA kernels (FP32 accumulation).  No reference
// counterpart: the taper design loop uses it to screen candidate geometries, and full FDTD
// verifies every design it proposes.
//
// Numerics (the oracle in tests/ follows the same quantisation points):
//   * master parameters FP32; GEMM operands (weights, activations, back-propagated errors) are
//     rounded to BF16 (round to nearest even); every GEMM accumulates in FP32 on the matrix cores;
//   * hidden pre-activations z are kept in FP32; h = bf16(GELU(z)) with GELU(z) = z/2 (1 + erf(z/sqrt 2));
//   * loss = mean over samples and outputs of (y - t)^2; Adam on the FP32 master parameters.
//
// GEMM: C(m, n) = sum_k A(m, k) B(k, n) with element strides for A and B, so one kernel serves the
// forward product X W^T, the input gradient dZ W and the weight gradient dZ^T X.  A block of four
// 64-wide wavefronts computes a 64 x 64 tile of C; each wavefront a 32 x 32 tile with
// v_mfma_f32_32x32x8_bf16 (lane l feeds A(m0 + l%32, k + 4(l/32) + r) and B(k + 4(l/32) + r,
// n0 + l%32), r = 0..3, and holds C(m0 + 8j + 4(l/32) + r, n0 + l%32) in accumulator 4j + r).
#include <cmath>
#include <cstring>
#include <vector>

#include "dev_internal.h"
#include "fdtd3d_mlp.h"

namespace fdtd {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__host__ __device__ __forceinline__ unsigned short to_bf16(float x) {
    unsigned int u = __builtin_bit_cast(unsigned int, x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float from_bf16(unsigned short h) { return __uint_as_float((unsigned int)h << 16); }

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

enum { EPI_STORE = 0, EPI_BIAS_GELU = 1, EPI_BIAS = 2 };

struct GemmArgs {
    const unsigned short *A;
    long long a_m, a_k;  // element strides of A(m, k)
    const unsigned short *B;
    long long b_k, b_n;  // element strides of B(k, n)
    int M, N, K;
    float *C;            // FP32 result (z for EPI_BIAS_GELU), row stride ldc
    long long ldc;
    const float *bias;   // per column n
    unsigned short *H;   // EPI_BIAS_GELU: bf16(GELU(z)), row stride ldh
    long long ldh;
};

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_bf16(GemmArgs g) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m0 = blockIdx.x * 64 + (w & 1) * 32, n0 = blockIdx.y * 64 + (w >> 1) * 32;
    const int row = m0 + (lane & 31), col = n0 + (lane & 31), kq = 4 * (lane >> 5);
    const bool mrow = row < g.M, ncol = col < g.N;
    f32x16 acc;
#pragma unroll
    for (int e = 0; e < 16; e++) acc[e] = 0.0f;
    for (int k = 0; k < g.K; k += 8) {
        bf16x4 a, b;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int kk = k + kq + r;
            const bool kin = kk < g.K;
            a[r] = (short)(mrow && kin ? g.A[row * g.a_m + kk * g.a_k] : 0);
            b[r] = (short)(ncol && kin ? g.B[kk * g.b_k + col * g.b_n] : 0);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    if (!ncol) return;
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int m = m0 + 8 * j + (lane >> 5) * 4 + r;
            if (m >= g.M) continue;
            float v = acc[4 * j + r];
            if (EPI != EPI_STORE) v += g.bias[col];
            g.C[m * g.ldc + col] = v;
            if (EPI == EPI_BIAS_GELU) g.H[m * g.ldh + col] = to_bf16(gelu(v));
        }
}

// Inputs (FP32 [n][n_in]) to bf16 rows padded to kin columns.
__global__ void k_pack_input(const float *x, int n, int n_in, int kin, unsigned short *out) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)n * kin) return;
    const int i = (int)(q / kin), c = (int)(q % kin);
    out[q] = c < n_in ? to_bf16(x[(long long)i * n_in + c]) : (unsigned short)0;
}

// dY = 2 (y - t) / (n n_out) (FP32 and its bf16 copy); loss partial sums per block.
__global__ void k_loss_grad(const float *y, const float *t, long long total, float scale, float *dz,
                            unsigned short *dzb, double *partial) {
    __shared__ double s[256];
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (q < total) {
        const float d = y[q] - t[q];
        e = (double)d * d;
        const float gq = scale * d;
        dz[q] = gq;
        dzb[q] = to_bf16(gq);
    }
    s[threadIdx.x] = e;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// dZ = dX * GELU'(z) for a hidden layer (FP32 and bf16 copies).
__global__ void k_gelu_back(const float *dx, const float *z, long long total, float *dz, unsigned short *dzb) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= total) return;
    const float v = dx[q] * gelu_grad(z[q]);
    dz[q] = v;
    dzb[q] = to_bf16(v);
}

// db[n] = sum over samples of dZ[m][n] (fixed order per column).
__global__ void k_bias_grad(const float *dz, int rows, int cols, float *db) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.0f;
    for (int m = 0; m < rows; m++) s += dz[(long long)m * cols + c];
    db[c] = s;
}

// Adam on the FP32 master parameters; refreshes the bf16 weight copies.
__global__ void k_adam(float *p, const float *g, float *mom, float *vel, long long n, float lr, float b1, float b2,
                       float eps, float c1, float c2, const long long *wseg, int nseg, unsigned short *wb) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float gq = g[q];
    const float m = b1 * mom[q] + (1.0f - b1) * gq;
    const float v = b2 * vel[q] + (1.0f - b2) * gq * gq;
    mom[q] = m;
    vel[q] = v;
    const float val = p[q] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    p[q] = val;
    // weight segments: [start, end) in the parameter vector -> offset in the bf16 weight buffer
    for (int s = 0; s < nseg; s++)
        if (q >= wseg[3 * s] && q < wseg[3 * s + 1]) wb[wseg[3 * s + 2] + (q - wseg[3 * s])] = to_bf16(val);
}

__global__ void k_to_bf16(const float *x, long long n, unsigned short *out) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = to_bf16(x[q]);
}

unsigned grid1(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace
}  // namespace fdtd

using namespace fdtd;

struct fdtd_mlp {
    int n_in, kin, width, depth, n_out;
    int nl;                          // layers = depth + 1
    std::vector<int> fin, fout;      // per layer (fin of layer 0 is kin)
    std::vector<long long> woff, boff;  // offsets in the parameter vector (weights [fout][fin_real], bias)
    std::vector<long long> wboff;    // offsets in the bf16 weight buffer ([fout][fin], fin padded)
    long long np = 0, nwb = 0;
    float *p = nullptr, *g = nullptr, *mom = nullptr, *vel = nullptr;
    unsigned short *wb = nullptr;
    long long *wseg = nullptr;
    long long step = 0;
    int cap = 0;                     // batch capacity of the activation buffers
    unsigned short *xin = nullptr;   // [cap][kin]
    std::vector<float *> z;          // per hidden layer [cap][width]; z[nl-1] = y [cap][n_out]
    std::vector<unsigned short *> h; // per hidden layer bf16 activations
    float *dz = nullptr, *dx = nullptr;
    unsigned short *dzb = nullptr;
    float *xf = nullptr, *tf = nullptr;
    double *partial = nullptr;
    hipStream_t stream = nullptr;
};

namespace {

// Layer l's weight as GEMM operand B(k, n) = W[n][k]; activation input A(m, k).
int gemm(fdtd_mlp *m, int epi, const unsigned short *A, long long a_m, long long a_k, const unsigned short *B,
         long long b_k, long long b_n, int M, int N, int K, float *C, long long ldc, const float *bias,
         unsigned short *H, long long ldh) {
    GemmArgs g{A, a_m, a_k, B, b_k, b_n, M, N, K, C, ldc, bias, H, ldh};
    dim3 grid((unsigned)((M + 63) / 64), (unsigned)((N + 63) / 64));
    if (epi == EPI_STORE) hipLaunchKernelGGL(k_gemm_bf16<EPI_STORE>, grid, dim3(256), 0, m->stream, g);
    else if (epi == EPI_BIAS) hipLaunchKernelGGL(k_gemm_bf16<EPI_BIAS>, grid, dim3(256), 0, m->stream, g);
    else hipLaunchKernelGGL(k_gemm_bf16<EPI_BIAS_GELU>, grid, dim3(256), 0, m->stream, g);
    FDTD_TRY(hipGetLastError());
    return 0;
}

void free_buffers(fdtd_mlp *m) {
    (void)dev_free(m->xin);
    for (auto *q : m->z) (void)dev_free(q);
    for (auto *q : m->h) (void)dev_free(q);
    (void)dev_free(m->dz);
    (void)dev_free(m->dx);
    (void)dev_free(m->dzb);
    (void)dev_free(m->xf);
    (void)dev_free(m->tf);
    (void)dev_free(m->partial);
    m->xin = nullptr;
    m->z.assign(m->nl, nullptr);
    m->h.assign(m->nl, nullptr);
    m->dz = m->dx = nullptr;
    m->dzb = nullptr;
    m->xf = m->tf = nullptr;
    m->partial = nullptr;
    m->cap = 0;
}

int ensure_capacity(fdtd_mlp *m, int n) {
    if (n <= m->cap) return 0;
    FDTD_TRY(hipStreamSynchronize(m->stream));
    free_buffers(m);
    const size_t N = (size_t)n;
    const int wmax = m->width > m->n_out ? m->width : m->n_out;
    FDTD_TRY(dev_malloc(&m->xin, N * m->kin * sizeof(unsigned short)));
    for (int l = 0; l < m->nl; l++) {
        FDTD_TRY(dev_malloc(&m->z[l], N * m->fout[l] * sizeof(float)));
        if (l < m->nl - 1) FDTD_TRY(dev_malloc(&m->h[l], N * m->fout[l] * sizeof(unsigned short)));
    }
    const int dmax = wmax > m->kin ? wmax : m->kin;
    FDTD_TRY(dev_malloc(&m->dz, N * dmax * sizeof(float)));
    FDTD_TRY(dev_malloc(&m->dx, N * dmax * sizeof(float)));
    FDTD_TRY(dev_malloc(&m->dzb, N * dmax * sizeof(unsigned short)));
    FDTD_TRY(dev_malloc(&m->xf, N * m->n_in * sizeof(float)));
    FDTD_TRY(dev_malloc(&m->tf, N * m->n_out * sizeof(float)));
    FDTD_TRY(dev_malloc(&m->partial, (N * m->n_out / 256 + 2) * sizeof(double)));
    m->cap = n;
    return 0;
}

// Forward pass of n samples already packed in m->xin; y lands in m->z[nl-1].
int forward(fdtd_mlp *m, int n) {
    const unsigned short *a = m->xin;
    long long lda = m->kin;
    for (int l = 0; l < m->nl; l++) {
        const bool last = l == m->nl - 1;
        int rc = gemm(m, last ? EPI_BIAS : EPI_BIAS_GELU, a, lda, 1, m->wb + m->wboff[l], 1, m->fin[l], n, m->fout[l],
                      m->fin[l], m->z[l], m->fout[l], m->p + m->boff[l], last ? nullptr : m->h[l], m->fout[l]);
        if (rc) return rc;
        if (!last) {
            a = m->h[l];
            lda = m->fout[l];
        }
    }
    return 0;
}

int upload_inputs(fdtd_mlp *m, const float *x, int n) {
    int rc = ensure_capacity(m, n);
    if (rc) return rc;
    FDTD_TRY(hipMemcpyAsync(m->xf, x, (size_t)n * m->n_in * sizeof(float), hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(k_pack_input, dim3(grid1((long long)n * m->kin)), dim3(256), 0, m->stream, (const float *)m->xf,
                       n, m->n_in, m->kin, m->xin);
    FDTD_TRY(hipGetLastError());
    return 0;
}

// Loss and the gradient of every parameter into m->g (FP32, parameter layout).
int backward(fdtd_mlp *m, const float *t, int n, double *loss) {
    const long long tot = (long long)n * m->n_out;
    FDTD_TRY(hipMemcpyAsync(m->tf, t, (size_t)tot * sizeof(float), hipMemcpyHostToDevice, m->stream));
    const unsigned nb = grid1(tot);
    hipLaunchKernelGGL(k_loss_grad, dim3(nb), dim3(256), 0, m->stream, (const float *)m->z[m->nl - 1],
                       (const float *)m->tf, tot, 2.0f / (float)tot, m->dz, m->dzb, m->partial);
    FDTD_TRY(hipGetLastError());
    for (int l = m->nl - 1; l >= 0; l--) {
        const int fo = m->fout[l], fi = m->fin[l];
        const unsigned short *x = l == 0 ? m->xin : m->h[l - 1];
        // dW[o][i] = sum_m dZ(m, o) X(m, i): A(o, m) = dzb[m*fo + o], B(m, i) = x[m*fi + i]
        int rc = gemm(m, EPI_STORE, m->dzb, 1, fo, x, fi, 1, fo, fi, n, m->g + m->woff[l], fi, nullptr, nullptr, 0);
        if (rc) return rc;
        hipLaunchKernelGGL(k_bias_grad, dim3(grid1(fo)), dim3(256), 0, m->stream, (const float *)m->dz, n, fo,
                           m->g + m->boff[l]);
        FDTD_TRY(hipGetLastError());
        if (l == 0) break;
        // dX(m, i) = sum_o dZ(m, o) W[o][i]: A(m, o) = dzb[m*fo + o], B(o, i) = wb[o*fi + i]
        rc = gemm(m, EPI_STORE, m->dzb, fo, 1, m->wb + m->wboff[l], fi, 1, n, fi, fo, m->dx, fi, nullptr, nullptr, 0);
        if (rc) return rc;
        hipLaunchKernelGGL(k_gelu_back, dim3(grid1((long long)n * fi)), dim3(256), 0, m->stream, (const float *)m->dx,
                           (const float *)m->z[l - 1], (long long)n * fi, m->dz, m->dzb);
        FDTD_TRY(hipGetLastError());
    }
    std::vector<double> part(nb);
    FDTD_TRY(hipMemcpyAsync(part.data(), m->partial, nb * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    FDTD_TRY(hipStreamSynchronize(m->stream));
    double s = 0.0;
    for (double v : part) s += v;
    if (loss) *loss = s / (double)tot;
    return 0;
}

}  // namespace

extern "C" {

fdtd_mlp *fdtd_mlp_new(int n_in, int width, int depth, int n_out) {
    if (n_in < 1 || width < 1 || depth < 1 || n_out < 1) {
        set_error("mlp: sizes must be positive");
        return nullptr;
    }
    fdtd_mlp *m = new fdtd_mlp();
    m->n_in = n_in;
    m->kin = (n_in + 7) / 8 * 8;
    m->width = width;
    m->depth = depth;
    m->n_out = n_out;
    m->nl = depth + 1;
    long long off = 0, woff = 0;
    for (int l = 0; l < m->nl; l++) {
        const int fi = l == 0 ? m->kin : width, fi_real = l == 0 ? n_in : width;
        const int fo = l == m->nl - 1 ? n_out : width;
        m->fin.push_back(fi);
        m->fout.push_back(fo);
        m->woff.push_back(off);
        off += (long long)fo * fi_real;
        m->boff.push_back(off);
        off += fo;
        m->wboff.push_back(woff);
        woff += (long long)fo * fi;
    }
    m->np = off;
    m->nwb = woff;
    m->z.assign(m->nl, nullptr);
    m->h.assign(m->nl, nullptr);
    // gradients use the padded layout for layer 0's weights too; the parameter vector keeps the
    // real input count, so layer 0's gradient GEMM writes into a padded scratch first (see
    // fdtd_mlp_loss_grad)
    hipError_t e = dev_stream_acquire(&m->stream);
    if (e == hipSuccess) e = dev_malloc(&m->p, (size_t)m->np * sizeof(float));
    if (e == hipSuccess) e = dev_malloc(&m->g, (size_t)(m->np + (long long)m->fout[0] * m->kin) * sizeof(float));
    if (e == hipSuccess) e = dev_malloc(&m->mom, (size_t)m->np * sizeof(float));
    if (e == hipSuccess) e = dev_malloc(&m->vel, (size_t)m->np * sizeof(float));
    if (e == hipSuccess) e = dev_malloc(&m->wb, (size_t)m->nwb * sizeof(unsigned short));
    if (e == hipSuccess) e = dev_malloc(&m->wseg, (size_t)3 * m->nl * sizeof(long long));
    if (e == hipSuccess) e = hipMemset(m->mom, 0, (size_t)m->np * sizeof(float));
    if (e == hipSuccess) e = hipMemset(m->vel, 0, (size_t)m->np * sizeof(float));
    if (e == hipSuccess) e = hipMemset(m->wb, 0, (size_t)m->nwb * sizeof(unsigned short));
    if (e != hipSuccess) {
        fail("mlp create", e);
        fdtd_mlp_free(m);
        return nullptr;
    }
    return m;
}

void fdtd_mlp_free(fdtd_mlp *m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    free_buffers(m);
    (void)dev_free(m->p);
    (void)dev_free(m->g);
    (void)dev_free(m->mom);
    (void)dev_free(m->vel);
    (void)dev_free(m->wb);
    (void)dev_free(m->wseg);
    if (m->stream) (void)dev_stream_release(m->stream);
    delete m;
}

long long fdtd_mlp_param_count(const fdtd_mlp *m) { return m->np; }

int fdtd_mlp_set_params(fdtd_mlp *m, const float *params) {
    FDTD_TRY(hipStreamSynchronize(m->stream));
    FDTD_TRY(hipMemcpy(m->p, params, (size_t)m->np * sizeof(float), hipMemcpyHostToDevice));
    // bf16 weights in the padded [fout][fin] layout (layer 0's padding columns stay zero)
    std::vector<unsigned short> wb((size_t)m->nwb, 0);
    std::vector<long long> seg;
    for (int l = 0; l < m->nl; l++) {
        const int fi_real = l == 0 ? m->n_in : m->width, fi = m->fin[l];
        for (int o = 0; o < m->fout[l]; o++)
            for (int i = 0; i < fi_real; i++)
                wb[(size_t)(m->wboff[l] + (long long)o * fi + i)] = to_bf16(params[m->woff[l] + (long long)o * fi_real + i]);
        if (fi == fi_real) {
            seg.push_back(m->woff[l]);
            seg.push_back(m->woff[l] + (long long)m->fout[l] * fi);
            seg.push_back(m->wboff[l]);
        }
    }
    FDTD_TRY(hipMemcpy(m->wb, wb.data(), wb.size() * sizeof(unsigned short), hipMemcpyHostToDevice));
    FDTD_TRY(hipMemcpy(m->wseg, seg.data(), seg.size() * sizeof(long long), hipMemcpyHostToDevice));
    FDTD_TRY(hipMemset(m->mom, 0, (size_t)m->np * sizeof(float)));
    FDTD_TRY(hipMemset(m->vel, 0, (size_t)m->np * sizeof(float)));
    m->step = 0;
    return 0;
}

int fdtd_mlp_get_params(fdtd_mlp *m, float *params) {
    FDTD_TRY(hipStreamSynchronize(m->stream));
    FDTD_TRY(hipMemcpy(params, m->p, (size_t)m->np * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int fdtd_mlp_forward(fdtd_mlp *m, const float *x, int n, float *y) {
    if (n <= 0) return 0;
    int rc = upload_inputs(m, x, n);
    if (rc || (rc = forward(m, n))) return rc;
    FDTD_TRY(hipMemcpyAsync(y, m->z[m->nl - 1], (size_t)n * m->n_out * sizeof(float), hipMemcpyDeviceToHost,
                            m->stream));
    FDTD_TRY(hipStreamSynchronize(m->stream));
    return 0;
}

// Layer 0's weight gradient is formed over the padded input width and compacted to the real one.
static int grads(fdtd_mlp *m, const float *x, const float *t, int n, double *loss) {
    int rc = upload_inputs(m, x, n);
    if (rc || (rc = forward(m, n))) return rc;
    // point layer 0's weight gradient at the padded scratch behind the parameter-sized buffer
    const long long w0 = m->woff[0];
    m->woff[0] = m->np;
    rc = backward(m, t, n, loss);
    m->woff[0] = w0;
    if (rc) return rc;
    std::vector<float> pad((size_t)m->fout[0] * m->kin);
    FDTD_TRY(hipMemcpy(pad.data(), m->g + m->np, pad.size() * sizeof(float), hipMemcpyDeviceToHost));
    std::vector<float> real((size_t)m->fout[0] * m->n_in);
    for (int o = 0; o < m->fout[0]; o++)
        for (int i = 0; i < m->n_in; i++) real[(size_t)o * m->n_in + i] = pad[(size_t)o * m->kin + i];
    FDTD_TRY(hipMemcpy(m->g + w0, real.data(), real.size() * sizeof(float), hipMemcpyHostToDevice));
    return 0;
}

int fdtd_mlp_loss_grad(fdtd_mlp *m, const float *x, const float *t, int n, double *loss, float *grad) {
    if (n <= 0) {
        set_error("mlp: empty batch");
        return -2;
    }
    int rc = grads(m, x, t, n, loss);
    if (rc) return rc;
    if (grad) FDTD_TRY(hipMemcpy(grad, m->g, (size_t)m->np * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int fdtd_mlp_adam_step(fdtd_mlp *m, const float *x, const float *t, int n, float lr, float beta1, float beta2,
                       float eps, double *loss) {
    if (n <= 0) {
        set_error("mlp: empty batch");
        return -2;
    }
    int rc = grads(m, x, t, n, loss);
    if (rc) return rc;
    m->step++;
    const float c1 = 1.0f - std::pow(beta1, (float)m->step), c2 = 1.0f - std::pow(beta2, (float)m->step);
    int nseg = 0;
    for (int l = 1; l < m->nl; l++) nseg++;  // layers with unpadded inputs (all but layer 0 when padded)
    if (m->kin == m->n_in) nseg++;
    hipLaunchKernelGGL(k_adam, dim3(grid1(m->np)), dim3(256), 0, m->stream, m->p, (const float *)m->g, m->mom, m->vel,
                       m->np, lr, beta1, beta2, eps, c1, c2, (const long long *)m->wseg, nseg, m->wb);
    FDTD_TRY(hipGetLastError());
    if (m->kin != m->n_in) {
        // layer 0's bf16 weights live in the padded layout: refresh them from the updated FP32 rows
        std::vector<float> w((size_t)m->fout[0] * m->n_in);
        FDTD_TRY(hipMemcpyAsync(w.data(), m->p + m->woff[0], w.size() * sizeof(float), hipMemcpyDeviceToHost,
                                m->stream));
        FDTD_TRY(hipStreamSynchronize(m->stream));
        std::vector<unsigned short> wb((size_t)m->fout[0] * m->kin, 0);
        for (int o = 0; o < m->fout[0]; o++)
            for (int i = 0; i < m->n_in; i++) wb[(size_t)o * m->kin + i] = to_bf16(w[(size_t)o * m->n_in + i]);
        FDTD_TRY(hipMemcpy(m->wb + m->wboff[0], wb.data(), wb.size() * sizeof(unsigned short), hipMemcpyHostToDevice));
    }
    return 0;
}

}  // extern "C"
