// Single-pass leapfrog launches (dev_leap.hip): one kernel per time step that reads every field
// component once and writes it once.  Shared by the device-resident runs (dev_run.hip) and the
// batched runs (dev_batch.hip).
#pragma once

#include "dev_internal.h"

namespace fdtd {

// One leap launch over one grid.  mode 1 advances H of step s and then E of step s + 1 (fields
// and CPML memories read from src / psrc, written to dst / pdst, the other buffer); mode 0 is the
// E update of step s alone and mode 2 the H update of step s alone (in place: dst == src, pdst ==
// psrc).  A run of n steps is one mode-0 launch, n - 1 mode-1 launches and one mode-2 launch: the
// updates of n split E / H pairs, in the same order.
//
// Compact on purpose: every buffer is one base pointer plus 32-bit offsets (component m of cell
// c at base[m * nal + c]; CPML memory (axis a, field f = 0 E / 1 H, component q) at
// psi[poff[a] + (2 f + q) * pn[a] + index]; factors of axis a at tab[toff[a] + ...]), so that the
// kernel's uniform state fits the scalar registers instead of spilling on every plane.
struct LeapArgs {
    const float *src;
    float *dst;
    int nal;
    const void *mat;
    const float *coef;  // [nmat][kCoefStride], semi-implicit scheme
    int nmat;           // materials copied into LDS (0: coefficients read from global memory)
    const float *tab;
    int toff[3];
    const float *psrc;
    float *pdst;
    int poff[3], pn[3];
    int nlo[3], nhi[3], W[3];
    int I, J, K, P;
    int mode;
    int tw, th;                 // owned columns / rows per tile
    int nwx, nwy, kc_len, nkc;  // tiles: x strips, y strips, z chunks
    // source planes (6 x P each: J, then M), static during the launch
    int nsrc;
    int src_k0, src_k1;         // planes [src_k0, src_k1) may carry one
    int srck[kMaxSrcPlanes];
    const float *srcJ[kMaxSrcPlanes];
    // overlap monitors folded into the launch that forms H(s): one partial per wavefront and step
    int nmon;
    RunMonArg mon[kMaxRunMon];
    const int *step;            // device step counter (batched runs), or null
    int step0;
    long long vb0;              // batched runs: first block of this grid
};

// Tile geometry: a block of kLeapBlockH wavefronts of kLeapBlockW lanes covers an extended tile of
// kLeapBlockW x kLeapBlockH cells; the inner (kLeapBlockW - 2) x (kLeapBlockH - 2) are owned.
constexpr int kLeapBlockW = 64, kLeapBlockH = 16;
constexpr int kLeapCoef = 18;      // coefficient floats per material kept in LDS: caE, cbE, cjE, caH, cbH, cjH x 3
constexpr int kLeapLdsMats = 512;  // largest table copied into LDS (36 KB)

// Host helpers (dev_leap.hip).
// True when grid d can run leap launches: PML boundary, no dense source volume, every buffer
// addressable with 32-bit offsets (6 x component stride below 2^30 floats).
bool leap_supported(const fdtd_dev *d);
// Monitor partial slots per step (one per wavefront of a plane's tiles).
int leap_plane_slots(const fdtd_dev *d);
// Fill L for grid d in `mode` with z chunks of kc_len planes; destination = the second buffer in
// mode 1, the same buffer otherwise; stored source planes from d; no monitors.  The coefficient
// and CPML tables for dt and fused_prepare(d) must be ready.
void leap_fill(fdtd_dev *d, int mode, int kc_len, LeapArgs *L);
// Blocks of one launch over L's grid.
long long leap_blocks(const LeapArgs &L);
// Dynamic LDS bytes of a launch over L's grid.
size_t leap_lds_bytes(const LeapArgs &L);
// Launch one leap over a single grid on stream st (events t0/t1 bracket it when non-null).
int leap_launch(const LeapArgs &L, bool u8, hipStream_t st, hipEvent_t t0, hipEvent_t t1);
// Default z planes per chunk for grid d.
int leap_default_chunk(const fdtd_dev *d);

}  // namespace fdtd
