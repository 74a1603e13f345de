// Single-pass full steps (dev_onepass.hip): one launch per time step advances E and then H of the
// same step over every interior cell of a CPML grid, reading each field component once and writing
// it once into the second field buffer.  Shared by the single-grid steps (dev_yee.hip, dev_run.hip)
// and the batched runs (dev_batch.hip).
#pragma once

#include "dev_stencil.h"

namespace fdtd {

// Launch state of one grid.  `a` holds the current buffers (a.Ex .. a.Hz), materials,
// coefficients, sources and monitors exactly as a split E / H pair would; the step writes fields
// and CPML memories into the second buffers (dst, pdst).  Byte offsets address every buffer from
// one base, so the whole field buffer and the whole CPML memory each stay below 2 GiB.
struct OnePassArgs {
    YeeArgs a;
    const float *coefp;   // packed semi-implicit coefficients, kCoefPacked floats per material
    const float *cbe;     // cbE per material (isotropic lossless tables)
    float cbh;            // the grid's cbH (isotropic lossless tables)
    int iso;              // the table is isotropic and lossless (the update reads cbe and cbh only)
    const float *tab[3];  // CPML factors per axis: bE, aE, kE, bH, aH, kH (0, 0, 1 outside the slabs)
    int nlo[3], nhi[3], W[3];
    const float *psrc;    // current CPML memories (one allocation)
    float *pdst;          // second CPML memories, same layout
    unsigned pbytes;      // bytes of one CPML allocation
    unsigned poE[3][2], poH[3][2];  // byte offset of each memory inside the allocation
    float *dst;           // second field buffer: component m at dst + m * nal
    unsigned nal4;        // component stride in bytes
    int ntx, nty, tiles;  // tiles per plane (x, y, product)
    int tw, th;           // owned columns and rows per tile
    int kc;               // z planes per chunk
    int nkc;              // z chunks
    // tiles clear of the x slabs [bx0, bx1), of the y slabs [by0, by1); chunks clear of the z slabs
    // [bk0, bk1); blocks per slab class: clear, x only, y only, x and y (clear chunks), z chunks
    int bx0, bx1, by0, by1, bk0, bk1;
    int ncls[5];
    long long vb0;        // batched runs: first block of this grid
};

// Threads per block of the single-pass step (four wavefronts).
constexpr int kOpBlock = 256;

// True when grid d can take single-pass steps: CPML boundary, 1- or 2-byte material ids, at most
// kMaxSrcPlanes stored source planes and no dense source volume, fields and CPML memories below
// 2 GiB each (byte offsets in 31 bits).
bool onepass_supported(const fdtd_dev *d);
// Tiles per z plane for x-tile width tw (64 or 32 columns): the monitor partial slots per step.
int onepass_plane_slots(const fdtd_dev *d, int tw);
// Default z planes per chunk and tile width for grid d.
int onepass_default_chunk(const fdtd_dev *d);
int onepass_default_width(const fdtd_dev *d);
// Fill the launch state of grid d for one step (sources and monitors from d's run state, as
// fill_step_args + the run fields of the split path).  fused_prepare(d) and the coefficient and
// CPML tables for dt must be ready.
int onepass_fill(fdtd_dev *d, double dt, int tw, int kc, OnePassArgs *A);
// One single-pass step of grid d (fields and memories ping-pong; fused_swap afterwards).
int onepass_step(fdtd_dev *d, double dt, int tw, int kc);

}  // namespace fdtd
