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
// GEMM: every product is put in the form C[M][N] = A[M][K] B[N][K]^T with K contiguous in both
// operands: the forward product X W^T directly, the input gradient dZ W with a transposed bf16 copy
// of W (kept current by the Adam kernel), the weight gradient dZ^T X from transposed copies of dZ and
// X (one LDS transpose each, batch dimension padded to a multiple of 8 with zeros).  A block of four
// 64-wide wavefronts computes a 128 x 128 tile of C: 128 x 32 slices of A and B go through LDS with
// 16-byte global loads (double-buffered: the next slice's loads are in flight during this slice's
// matrix-core work), and each wavefront owns a 64 x 64 sub-tile, 2 x 2 tiles of gfx950's
// v_mfma_f32_32x32x16_bf16 (lane l feeds A(m0 + l%32, k + 8(l/32) + r) and B(k + 8(l/32) + r,
// n0 + l%32), r = 0..7, one 16-byte LDS read per operand, and holds C(m0 + 8j + 4(l/32) + r,
// n0 + l%32), r = 0..3, in accumulator 4j + r): K = 16 per instruction, twice the rate of the
// 32x32x8 form.
#include <cmath>
#include <cstring>
#include <vector>

#include "dev_internal.h"
#include "fdtd3d_mlp.h"

namespace fdtd {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__host__ __device__ __forceinline__ unsigned short to_bf16(float x) {
    unsigned int u = __builtin_bit_cast(unsigned int, x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// EPI_GELU_ONLY: inference hidden layers, only h = bf16(GELU(z)) is written (z is needed only by
// the backward pass)
enum { EPI_STORE = 0, EPI_BIAS_GELU = 1, EPI_BIAS = 2, EPI_GELU_ONLY = 3 };

struct NtArgs {
    const unsigned short *A;  // [M][lda], K contiguous
    const unsigned short *B;  // [N][ldb], K contiguous
    long long lda, ldb;       // multiples of 8 elements
    int M, N, K;              // K a multiple of 8 (operands zero-padded up to it)
    float *C;                 // FP32 result (z for EPI_BIAS_GELU), row stride ldc
    long long ldc;
    int kchunk;               // split K: block z sums k in [z kchunk, (z + 1) kchunk) into C + z cz
    long long cz;
    const float *bias;        // per column n
    unsigned short *H;        // EPI_BIAS_GELU: bf16(GELU(z)), row stride ldh
    long long ldh;
};

constexpr int kGBM = 128, kGBN = 128, kGBK = 32;
constexpr int kSplitK = 128, kMaxSplit = 64;  // split-K chunk target and limit
constexpr int kGLD = kGBK + 8;  // LDS row stride (elements): 80 B keeps 16-byte stores aligned

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_nt(NtArgs g) {
    __shared__ __attribute__((aligned(16))) unsigned short As[2][kGBM * kGLD];
    __shared__ __attribute__((aligned(16))) unsigned short Bs[2][kGBN * kGLD];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int bm = blockIdx.x * kGBM, bn = blockIdx.y * kGBN;
    const int wm = (w & 1) * 64, wn = (w >> 1) * 64;
    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[i][j][e] = 0.0f;
    // each thread moves two 8-element chunks of A and two of B per K slice: chunk c = t + 256 q
    // covers row c / 4, elements 8 (c % 4) .. +7 of the slice
    uint4 ra[2], rb[2];
    const int kbeg = blockIdx.z * g.kchunk, kend = min(g.K, kbeg + g.kchunk);
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int c = t + 256 * q, row = c >> 2, k = k0 + (c & 3) * 8;
            ra[q] = make_uint4(0u, 0u, 0u, 0u);
            rb[q] = make_uint4(0u, 0u, 0u, 0u);
            if (bm + row < g.M && k < kend) ra[q] = *reinterpret_cast<const uint4 *>(g.A + (bm + row) * g.lda + k);
            if (bn + row < g.N && k < kend) rb[q] = *reinterpret_cast<const uint4 *>(g.B + (bn + row) * g.ldb + k);
        }
    };
    auto stash = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int c = t + 256 * q, row = c >> 2, kc = (c & 3) * 8;
            *reinterpret_cast<uint4 *>(&As[buf][row * kGLD + kc]) = ra[q];
            *reinterpret_cast<uint4 *>(&Bs[buf][row * kGLD + kc]) = rb[q];
        }
    };
    fetch(kbeg);
    stash(0);
    __syncthreads();
    const int nk = (kend - kbeg + kGBK - 1) / kGBK;
    const int fr = lane & 31, fk = 8 * (lane >> 5);
    for (int kt = 0; kt < nk; kt++) {
        const int cur = kt & 1;
        if (kt + 1 < nk) fetch(kbeg + (kt + 1) * kGBK);
#pragma unroll
        for (int ks = 0; ks < kGBK; ks += 16) {
            bf16x8 af[2], bf[2];
#pragma unroll
            for (int i = 0; i < 2; i++) {
                af[i] = *reinterpret_cast<const bf16x8 *>(&As[cur][(wm + 32 * i + fr) * kGLD + ks + fk]);
                bf[i] = *reinterpret_cast<const bf16x8 *>(&Bs[cur][(wn + 32 * i + fr) * kGLD + ks + fk]);
            }
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
        // the other buffer was last read in the previous slice, before its closing barrier
        if (kt + 1 < nk) stash(cur ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int col = bn + wn + 32 * j + fr;
        if (col >= g.N) continue;
        const float b = EPI != EPI_STORE ? g.bias[col] : 0.0f;
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int q = 0; q < 4; q++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int m = bm + wm + 32 * i + 8 * q + 4 * (lane >> 5) + r;
                    if (m >= g.M) continue;
                    const float v = acc[i][j][4 * q + r] + b;
                    if (EPI != EPI_GELU_ONLY) g.C[blockIdx.z * g.cz + m * g.ldc + col] = v;
                    if (EPI == EPI_BIAS_GELU || EPI == EPI_GELU_ONLY) g.H[m * g.ldh + col] = to_bf16(gelu(v));
                }
    }
}

// out[c][r] = in[r][c] for a bf16 [rows][cols] matrix (row stride ld_in) into [cols][ld_out], rows
// rows..rows_pad-1 of the result zero (64 x 64 tiles through LDS).
// Each thread moves groups of 4 elements (8-byte loads along a row, 8-byte stores along an output
// row) where the group lies inside the matrix and is 8-byte aligned, single elements otherwise.
__global__ __launch_bounds__(256) void k_transpose_bf16(const unsigned short *in, int rows, int cols, long long ld_in,
                                                        unsigned short *out, long long ld_out, int rows_pad) {
    __shared__ unsigned short tile[64][68];
    const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
    for (int q = threadIdx.x; q < 64 * 16; q += 256) {
        const int r = q >> 4, c = (q & 15) * 4;
        const unsigned short *src = in + (long long)(r0 + r) * ld_in + c0 + c;
        if (r0 + r < rows && c0 + c + 3 < cols && ((uintptr_t)src & 7) == 0) {
            const uint2 u = *reinterpret_cast<const uint2 *>(src);
            tile[r][c] = (unsigned short)(u.x & 0xffffu);
            tile[r][c + 1] = (unsigned short)(u.x >> 16);
            tile[r][c + 2] = (unsigned short)(u.y & 0xffffu);
            tile[r][c + 3] = (unsigned short)(u.y >> 16);
        } else {
            for (int e = 0; e < 4; e++)
                tile[r][c + e] = r0 + r < rows && c0 + c + e < cols ? src[e] : (unsigned short)0;
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < 64 * 16; q += 256) {
        const int c = q >> 4, r = (q & 15) * 4;
        if (c0 + c >= cols) continue;
        unsigned short *dst = out + (long long)(c0 + c) * ld_out + r0 + r;
        if (r0 + r + 3 < rows_pad && ((uintptr_t)dst & 7) == 0) {
            uint2 u;
            u.x = (unsigned)tile[r][c] | ((unsigned)tile[r + 1][c] << 16);
            u.y = (unsigned)tile[r + 2][c] | ((unsigned)tile[r + 3][c] << 16);
            *reinterpret_cast<uint2 *>(dst) = u;
        } else {
            for (int e = 0; e < 4; e++)
                if (r0 + r + e < rows_pad) dst[e] = tile[r + e][c];
        }
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
// Column sums of dz [rows][cols] (bias gradients) in a fixed order: block (x, y) sums rows
// kColRows y .. kColRows (y + 1) - 1 of columns 64 x .. 64 x + 63 (4 row lanes per column, then the
// 4 lanes in order) into part[y][col]; k_colsum_final adds the row blocks in order.  32 rows per
// block: a 4096-sample batch spreads over 512 blocks instead of idling most of the chip.
constexpr int kColRows = 32;
__global__ __launch_bounds__(256) void k_colsum_partial(const float *dz, int rows, int cols, float *part) {
    __shared__ float red[4][64];
    const int cl = threadIdx.x & 63, rl = threadIdx.x >> 6, col = blockIdx.x * 64 + cl;
    float s = 0.0f;
    if (col < cols)
        for (int r = blockIdx.y * kColRows + rl; r < rows && r < (blockIdx.y + 1) * kColRows; r += 4)
            s += dz[(long long)r * cols + col];
    red[rl][cl] = s;
    __syncthreads();
    if (rl == 0 && col < cols) part[(long long)blockIdx.y * cols + col] = ((red[0][cl] + red[1][cl]) + red[2][cl]) + red[3][cl];
}

// Block of 256 threads = 32 columns x 8 lanes: lane l sums row blocks l, l + 8, ... in order, then
// the 8 lane sums are added pairwise in a fixed tree.
__global__ __launch_bounds__(256) void k_colsum_final(const float *part, int nblk, int cols, float *db) {
    __shared__ float red[8][32];
    const int cl = threadIdx.x & 31, l = threadIdx.x >> 5, c = blockIdx.x * 32 + cl;
    float s = 0.0f;
    if (c < cols)
        for (int b = l; b < nblk; b += 8) s += part[(long long)b * cols + c];
    red[l][cl] = s;
    __syncthreads();
    if (l == 0 && c < cols)
        db[c] = ((red[0][cl] + red[1][cl]) + (red[2][cl] + red[3][cl])) +
                ((red[4][cl] + red[5][cl]) + (red[6][cl] + red[7][cl]));
}

// C[i] = sum over z (in order) of ws[z * mn + i]: the split-K partial products.
__global__ void k_splitk_reduce(const float *ws, int splits, long long mn, float *C) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= mn) return;
    float s = 0.0f;
    for (int z = 0; z < splits; z++) s += ws[z * mn + i];
    C[i] = s;
}

// Adam on the FP32 master parameters; refreshes the bf16 weight copies.
__global__ void k_adam(float *p, const float *g, float *mom, float *vel, long long n, float lr, float b1, float b2,
                       float eps, float c1, float c2, const long long *wseg, int nseg, unsigned short *wb,
                       unsigned short *wbt) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float gq = g[q];
    const float m = b1 * mom[q] + (1.0f - b1) * gq;
    const float v = b2 * vel[q] + (1.0f - b2) * gq * gq;
    mom[q] = m;
    vel[q] = v;
    const float val = p[q] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    p[q] = val;
    // weight segments (5 entries each): [start, end) in the parameter vector, offset in the bf16
    // weight buffer, offset of the transposed copy (-1: none), input width fin
    for (int s = 0; s < nseg; s++) {
        const long long *e = wseg + 5 * s;
        if (q < e[0] || q >= e[1]) continue;
        const long long rel = q - e[0];
        const unsigned short h = to_bf16(val);
        wb[e[2] + rel] = h;
        if (e[3] >= 0) {
            const long long fi = e[4], fo = (e[1] - e[0]) / fi, o = rel / fi, i = rel - o * fi;
            wbt[e[3] + i * fo + o] = h;
        }
    }
}

// Layer 0 with a padded input width: its weight-gradient rows [rows][src_cols] -> [rows][dst_cols]
// (the parameter layout keeps the real input count).
__global__ void k_compact_rows(const float *src, int rows, int src_cols, int dst_cols, float *dst) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)rows * dst_cols) return;
    const long long r = q / dst_cols, c = q - r * dst_cols;
    dst[q] = src[r * src_cols + c];
}

// Layer 0's bf16 weights in the padded layout from its FP32 rows; padding columns are zero.
__global__ void k_pad_rows_bf16(const float *src, int rows, int src_cols, int dst_cols, unsigned short *dst) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)rows * dst_cols) return;
    const long long r = q / dst_cols, c = q - r * dst_cols;
    dst[q] = c < src_cols ? to_bf16(src[r * src_cols + c]) : (unsigned short)0;
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
    std::vector<long long> wbtoff;   // offsets of the transposed copies ([fin][fout]) of layers 1.. in wbt
    long long nwbt = 0;
    unsigned short *wbt = nullptr;
    long long *wseg = nullptr;
    long long step = 0;
    int cap = 0;                     // batch capacity of the activation buffers
    unsigned short *xin = nullptr;   // [cap][kin]
    std::vector<float *> z;          // per hidden layer [cap][width]; z[nl-1] = y [cap][n_out]
    std::vector<unsigned short *> h; // per hidden layer bf16 activations
    float *dz = nullptr, *dx = nullptr;
    unsigned short *dzb = nullptr;
    unsigned short *dzt = nullptr, *xt = nullptr;  // transposed dZ and layer inputs, [features][npad]
    int npad = 0;                                  // batch capacity rounded up to a multiple of 8
    float *ws = nullptr;                           // split-K partial products, kMaxSplit x M x N
    float *colpart = nullptr;                      // bias-gradient row-block partials
    float *xf = nullptr, *tf = nullptr;
    double *partial = nullptr;
    hipStream_t stream = nullptr;
    // GEMM timing: when on, every GEMM launch records a start/stop event pair at its dispatch
    bool time_gemm = false;
    std::vector<hipEvent_t> gev;
    double gemm_flop = 0.0;  // multiply-adds x 2 of the timed launches
    hipEvent_t ev0 = nullptr, ev1 = nullptr;  // stream timer
};

namespace {

// C[M][N] = A[M][K] B[N][K]^T (+ bias, GELU) on the network's stream (see k_gemm_nt).  With few
// output tiles and a long K (the weight gradients: K = batch), K is split over up to kMaxSplit
// blocks per tile whose partial products are added in a fixed order (C then dense, ldc = N).
int gemm(fdtd_mlp *m, int epi, const unsigned short *A, long long lda, const unsigned short *B, long long ldb, int M,
         int N, int K, float *C, long long ldc, const float *bias, unsigned short *H, long long ldh) {
    if (lda % 8 || ldb % 8 || K % 8) {
        set_error("mlp gemm: row strides and K must be multiples of 8");
        return -2;
    }
    const int tiles = ((M + kGBM - 1) / kGBM) * ((N + kGBN - 1) / kGBN);
    int splits = 1;
    if (epi == EPI_STORE && ldc == N && tiles < 64 && K >= 2 * kSplitK) {
        splits = (K + kSplitK - 1) / kSplitK;
        if (splits > kMaxSplit) splits = kMaxSplit;
    }
    const int kchunk = splits == 1 ? K : ((K + splits - 1) / splits + kGBK - 1) / kGBK * kGBK;
    splits = (K + kchunk - 1) / kchunk;
    NtArgs g{A, B, lda, ldb, M, N, K, splits == 1 ? C : m->ws, ldc, kchunk, (long long)M * N, bias, H, ldh};
    dim3 grid((unsigned)((M + kGBM - 1) / kGBM), (unsigned)((N + kGBN - 1) / kGBN), (unsigned)splits);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (m->time_gemm) {
        FDTD_TRY(hipEventCreate(&e0));
        FDTD_TRY(hipEventCreate(&e1));
        m->gev.push_back(e0);
        m->gev.push_back(e1);
        m->gemm_flop += 2.0 * M * N * K;
    }
    if (epi == EPI_STORE) hipExtLaunchKernelGGL(k_gemm_nt<EPI_STORE>, grid, dim3(256), 0, m->stream, e0, e1, 0, g);
    else if (epi == EPI_BIAS) hipExtLaunchKernelGGL(k_gemm_nt<EPI_BIAS>, grid, dim3(256), 0, m->stream, e0, e1, 0, g);
    else if (epi == EPI_GELU_ONLY)
        hipExtLaunchKernelGGL(k_gemm_nt<EPI_GELU_ONLY>, grid, dim3(256), 0, m->stream, e0, e1, 0, g);
    else hipExtLaunchKernelGGL(k_gemm_nt<EPI_BIAS_GELU>, grid, dim3(256), 0, m->stream, e0, e1, 0, g);
    FDTD_TRY(hipGetLastError());
    if (splits > 1) {
        const long long mn = (long long)M * N;
        hipLaunchKernelGGL(k_splitk_reduce, dim3(grid1(mn)), dim3(256), 0, m->stream, (const float *)m->ws, splits, mn, C);
        FDTD_TRY(hipGetLastError());
    }
    return 0;
}

// [rows][cols] bf16 (row stride ld_in) -> [cols][npad], zero beyond rows.
int transpose(fdtd_mlp *m, const unsigned short *in, int rows, int cols, long long ld_in, unsigned short *out) {
    dim3 grid((unsigned)((m->npad + 63) / 64), (unsigned)((cols + 63) / 64));
    hipLaunchKernelGGL(k_transpose_bf16, grid, dim3(256), 0, m->stream, in, rows, cols, ld_in, out, (long long)m->npad,
                       m->npad);
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
    (void)dev_free(m->dzt);
    (void)dev_free(m->xt);
    (void)dev_free(m->ws);
    (void)dev_free(m->colpart);
    m->ws = m->colpart = nullptr;
    (void)dev_free(m->xf);
    (void)dev_free(m->tf);
    (void)dev_free(m->partial);
    m->xin = nullptr;
    m->z.assign(m->nl, nullptr);
    m->h.assign(m->nl, nullptr);
    m->dz = m->dx = nullptr;
    m->dzb = m->dzt = m->xt = nullptr;
    m->npad = 0;
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
    m->npad = (n + 7) / 8 * 8;
    FDTD_TRY(dev_malloc(&m->dzt, (size_t)m->npad * wmax * sizeof(unsigned short)));
    FDTD_TRY(dev_malloc(&m->xt, (size_t)m->npad * (m->kin > m->width ? m->kin : m->width) * sizeof(unsigned short)));
    const size_t fmax = (size_t)(m->kin > wmax ? m->kin : wmax);
    FDTD_TRY(dev_malloc(&m->ws, (size_t)kMaxSplit * wmax * fmax * sizeof(float)));
    FDTD_TRY(dev_malloc(&m->colpart, (size_t)((n + kColRows - 1) / kColRows) * wmax * sizeof(float)));
    FDTD_TRY(dev_malloc(&m->xf, N * m->n_in * sizeof(float)));
    FDTD_TRY(dev_malloc(&m->tf, N * m->n_out * sizeof(float)));
    FDTD_TRY(dev_malloc(&m->partial, (N * m->n_out / 256 + 2) * sizeof(double)));
    m->cap = n;
    return 0;
}

// Forward pass of n samples already packed in m->xin; y lands in m->z[nl-1].  `train` keeps the
// hidden pre-activations z for the backward pass.
int forward(fdtd_mlp *m, int n, bool train) {
    const unsigned short *a = m->xin;
    long long lda = m->kin;
    for (int l = 0; l < m->nl; l++) {
        const bool last = l == m->nl - 1;
        int rc = gemm(m, last ? EPI_BIAS : (train ? EPI_BIAS_GELU : EPI_GELU_ONLY), a, lda, m->wb + m->wboff[l], m->fin[l], n, m->fout[l],
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

// Loss and the gradient of every parameter into m->g (FP32, parameter layout) for the targets
// already in m->tf; the loss is read back (one synchronisation) only when `loss` is given.
int backward(fdtd_mlp *m, int n, double *loss) {
    const long long tot = (long long)n * m->n_out;
    const unsigned nb = grid1(tot);
    hipLaunchKernelGGL(k_loss_grad, dim3(nb), dim3(256), 0, m->stream, (const float *)m->z[m->nl - 1],
                       (const float *)m->tf, tot, 2.0f / (float)tot, m->dz, m->dzb, m->partial);
    FDTD_TRY(hipGetLastError());
    for (int l = m->nl - 1; l >= 0; l--) {
        const int fo = m->fout[l], fi = m->fin[l];
        const unsigned short *x = l == 0 ? m->xin : m->h[l - 1];
        // dW[o][i] = sum_m dZ[m][o] X[m][i] = (dZ^T)(X^T)^T, both transposed over the padded batch
        int rc = transpose(m, m->dzb, n, fo, fo, m->dzt);
        if (rc || (rc = transpose(m, x, n, fi, fi, m->xt))) return rc;
        rc = gemm(m, EPI_STORE, m->dzt, m->npad, m->xt, m->npad, fo, fi, m->npad, m->g + m->woff[l], fi, nullptr,
                  nullptr, 0);
        if (rc) return rc;
        const int nblk = (n + kColRows - 1) / kColRows;
        hipLaunchKernelGGL(k_colsum_partial, dim3((unsigned)((fo + 63) / 64), (unsigned)nblk), dim3(256), 0, m->stream,
                           (const float *)m->dz, n, fo, m->colpart);
        hipLaunchKernelGGL(k_colsum_final, dim3((unsigned)((fo + 31) / 32)), dim3(256), 0, m->stream,
                           (const float *)m->colpart, nblk, fo,
                           m->g + m->boff[l]);
        FDTD_TRY(hipGetLastError());
        if (l == 0) break;
        // dX[m][i] = sum_o dZ[m][o] W[o][i] = dZ (W^T)^T with the transposed weight copy [fi][fo]
        rc = gemm(m, EPI_STORE, m->dzb, fo, m->wbt + m->wbtoff[l], fo, n, fi, fo, m->dx, fi, nullptr, nullptr, 0);
        if (rc) return rc;
        hipLaunchKernelGGL(k_gelu_back, dim3(grid1((long long)n * fi)), dim3(256), 0, m->stream, (const float *)m->dx,
                           (const float *)m->z[l - 1], (long long)n * fi, m->dz, m->dzb);
        FDTD_TRY(hipGetLastError());
    }
    if (!loss) return 0;
    std::vector<double> part(nb);
    FDTD_TRY(hipMemcpyAsync(part.data(), m->partial, nb * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    FDTD_TRY(hipStreamSynchronize(m->stream));
    double s = 0.0;
    for (double v : part) s += v;
    *loss = s / (double)tot;
    return 0;
}

int upload_targets(fdtd_mlp *m, const float *t, int n) {
    FDTD_TRY(hipMemcpyAsync(m->tf, t, (size_t)n * m->n_out * sizeof(float), hipMemcpyHostToDevice, m->stream));
    return 0;
}

}  // namespace

extern "C" {

fdtd_mlp *fdtd_mlp_new(int n_in, int width, int depth, int n_out) {
    if (n_in < 1 || width < 1 || depth < 1 || n_out < 1) {
        set_error("mlp: sizes must be positive");
        return nullptr;
    }
    if (width % 8 || n_out % 8) {
        set_error("mlp: width and n_out must be multiples of 8 (16-byte operand rows)");
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
        m->wbtoff.push_back(l == 0 ? -1 : m->nwbt);
        if (l > 0) m->nwbt += (long long)fo * fi;
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
    if (e == hipSuccess) e = dev_malloc(&m->wbt, (size_t)(m->nwbt > 0 ? m->nwbt : 1) * sizeof(unsigned short));
    if (e == hipSuccess) e = dev_malloc(&m->wseg, (size_t)5 * m->nl * sizeof(long long));
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
    (void)dev_free(m->wbt);
    (void)dev_free(m->wseg);
    for (auto e : m->gev) (void)hipEventDestroy(e);
    if (m->ev0) (void)hipEventDestroy(m->ev0);
    if (m->ev1) (void)hipEventDestroy(m->ev1);
    if (m->stream) (void)dev_stream_release(m->stream);
    delete m;
}

long long fdtd_mlp_param_count(const fdtd_mlp *m) { return m->np; }

int fdtd_mlp_set_params(fdtd_mlp *m, const float *params) {
    FDTD_TRY(hipStreamSynchronize(m->stream));
    FDTD_TRY(hipMemcpy(m->p, params, (size_t)m->np * sizeof(float), hipMemcpyHostToDevice));
    // bf16 weights in the padded [fout][fin] layout (layer 0's padding columns stay zero) and the
    // transposed [fin][fout] copies of layers 1..
    std::vector<unsigned short> wb((size_t)m->nwb, 0), wbt((size_t)(m->nwbt > 0 ? m->nwbt : 1), 0);
    std::vector<long long> seg;
    for (int l = 0; l < m->nl; l++) {
        const int fi_real = l == 0 ? m->n_in : m->width, fi = m->fin[l], fo = m->fout[l];
        for (int o = 0; o < fo; o++)
            for (int i = 0; i < fi_real; i++) {
                const unsigned short h = to_bf16(params[m->woff[l] + (long long)o * fi_real + i]);
                wb[(size_t)(m->wboff[l] + (long long)o * fi + i)] = h;
                if (l > 0) wbt[(size_t)(m->wbtoff[l] + (long long)i * fo + o)] = h;
            }
        if (fi == fi_real) {
            seg.push_back(m->woff[l]);
            seg.push_back(m->woff[l] + (long long)fo * fi);
            seg.push_back(m->wboff[l]);
            seg.push_back(m->wbtoff[l]);
            seg.push_back(fi);
        }
    }
    FDTD_TRY(hipMemcpy(m->wb, wb.data(), wb.size() * sizeof(unsigned short), hipMemcpyHostToDevice));
    FDTD_TRY(hipMemcpy(m->wbt, wbt.data(), wbt.size() * sizeof(unsigned short), hipMemcpyHostToDevice));
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
    if (rc || (rc = forward(m, n, false))) return rc;
    FDTD_TRY(hipMemcpyAsync(y, m->z[m->nl - 1], (size_t)n * m->n_out * sizeof(float), hipMemcpyDeviceToHost,
                            m->stream));
    FDTD_TRY(hipStreamSynchronize(m->stream));
    return 0;
}

// Gradient of the samples already on the device (m->xin, m->tf).  Layer 0's weight gradient is
// formed over the padded input width behind the parameter-sized buffer and compacted on the device.
static int grads_resident(fdtd_mlp *m, int n, double *loss) {
    int rc = forward(m, n, true);
    if (rc) return rc;
    const long long w0 = m->woff[0];
    if (m->kin != m->n_in) m->woff[0] = m->np;
    rc = backward(m, n, loss);
    m->woff[0] = w0;
    if (rc) return rc;
    if (m->kin != m->n_in) {
        const long long cnt = (long long)m->fout[0] * m->n_in;
        hipLaunchKernelGGL(k_compact_rows, dim3(grid1(cnt)), dim3(256), 0, m->stream, (const float *)(m->g + m->np),
                           m->fout[0], m->kin, m->n_in, m->g + w0);
        FDTD_TRY(hipGetLastError());
    }
    return 0;
}

static int grads(fdtd_mlp *m, const float *x, const float *t, int n, double *loss) {
    int rc = upload_inputs(m, x, n);
    if (rc || (rc = upload_targets(m, t, n))) return rc;
    return grads_resident(m, n, loss);
}

// One Adam update from the gradient in m->g; the bf16 weight copies follow the FP32 parameters.
static int adam_update(fdtd_mlp *m, float lr, float beta1, float beta2, float eps) {
    m->step++;
    const float c1 = 1.0f - std::pow(beta1, (float)m->step), c2 = 1.0f - std::pow(beta2, (float)m->step);
    int nseg = 0;
    for (int l = 1; l < m->nl; l++) nseg++;  // layers with unpadded inputs (all but layer 0 when padded)
    if (m->kin == m->n_in) nseg++;
    hipLaunchKernelGGL(k_adam, dim3(grid1(m->np)), dim3(256), 0, m->stream, m->p, (const float *)m->g, m->mom, m->vel,
                       m->np, lr, beta1, beta2, eps, c1, c2, (const long long *)m->wseg, nseg, m->wb, m->wbt);
    FDTD_TRY(hipGetLastError());
    if (m->kin != m->n_in) {
        const long long cnt = (long long)m->fout[0] * m->kin;
        hipLaunchKernelGGL(k_pad_rows_bf16, dim3(grid1(cnt)), dim3(256), 0, m->stream, (const float *)(m->p + m->woff[0]),
                           m->fout[0], m->n_in, m->kin, m->wb + m->wboff[0]);
        FDTD_TRY(hipGetLastError());
    }
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
    return adam_update(m, lr, beta1, beta2, eps);
}

int fdtd_mlp_train(fdtd_mlp *m, const float *x, const float *t, int n, int steps, float lr, float beta1, float beta2,
                   float eps, double *loss_first, double *loss_last) {
    if (n <= 0 || steps <= 0) {
        set_error("mlp: empty batch or no steps");
        return -2;
    }
    int rc = upload_inputs(m, x, n);
    if (rc || (rc = upload_targets(m, t, n))) return rc;
    for (int s = 0; s < steps; s++) {
        double *want = s == 0 ? loss_first : nullptr;
        if ((rc = grads_resident(m, n, want))) return rc;
        if ((rc = adam_update(m, lr, beta1, beta2, eps))) return rc;
    }
    if (loss_last) {  // the loss after the last update
        if ((rc = forward(m, n, false))) return rc;
        const long long tot = (long long)n * m->n_out;
        hipLaunchKernelGGL(k_loss_grad, dim3(grid1(tot)), dim3(256), 0, m->stream, (const float *)m->z[m->nl - 1],
                           (const float *)m->tf, tot, 2.0f / (float)tot, m->dz, m->dzb, m->partial);
        FDTD_TRY(hipGetLastError());
        std::vector<double> part(grid1(tot));
        FDTD_TRY(hipMemcpyAsync(part.data(), m->partial, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                                m->stream));
        FDTD_TRY(hipStreamSynchronize(m->stream));
        double sum = 0.0;
        for (double v : part) sum += v;
        *loss_last = sum / (double)tot;
    }
    FDTD_TRY(hipStreamSynchronize(m->stream));
    return 0;
}

int fdtd_mlp_gemm_timing(fdtd_mlp *m, int on) {
    FDTD_TRY(hipStreamSynchronize(m->stream));
    for (auto e : m->gev) (void)hipEventDestroy(e);
    m->gev.clear();
    m->gemm_flop = 0.0;
    m->time_gemm = on != 0;
    return 0;
}

int fdtd_mlp_gemm_times(fdtd_mlp *m, int *launches, double *ms, double *flop) {
    FDTD_TRY(hipStreamSynchronize(m->stream));
    double total = 0.0;
    for (size_t q = 0; q + 1 < m->gev.size(); q += 2) {
        float t = 0.0f;
        FDTD_TRY(hipEventElapsedTime(&t, m->gev[q], m->gev[q + 1]));
        total += t;
    }
    if (launches) *launches = (int)(m->gev.size() / 2);
    if (ms) *ms = total;
    if (flop) *flop = m->gemm_flop;
    for (auto e : m->gev) (void)hipEventDestroy(e);
    m->gev.clear();
    m->gemm_flop = 0.0;
    return 0;
}

int fdtd_mlp_timer_start(fdtd_mlp *m) {
    if (!m->ev0) FDTD_TRY(hipEventCreate(&m->ev0));
    if (!m->ev1) FDTD_TRY(hipEventCreate(&m->ev1));
    FDTD_TRY(hipEventRecord(m->ev0, m->stream));
    return 0;
}

int fdtd_mlp_timer_stop(fdtd_mlp *m, double *elapsed_ms) {
    FDTD_TRY(hipEventRecord(m->ev1, m->stream));
    FDTD_TRY(hipEventSynchronize(m->ev1));
    float t = 0.0f;
    FDTD_TRY(hipEventElapsedTime(&t, m->ev0, m->ev1));
    *elapsed_ms = t;
    return 0;
}

}  // extern "C"
