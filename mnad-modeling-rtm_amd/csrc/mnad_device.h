// mnad_device.h -- device-side descriptors shared by the HIP kernels and the
// host runtime (plain structs, no torch types).
//
// Grid storage (per field, per shot): rows of `pitch` floats; grid point
// (j, i) (j = z row, i = x column, 0-based, padded grid incl. CPML) lives at
//     (j + MNAD_ZO) * pitch + (i + MNAD_XO).
// Two zero rows above, 2 + MNAD_ZPAD below and at least two zero columns
// left/right are never written, so the +-2 stencil reads zeros beyond the grid, which is what
// the *_gen.m index ranges do (neighbours beyond the padded grid omitted).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "mnad_ring.h"
#include "mnad_stencil.h"

#define MNAD_ZO 2    // ghost rows above the grid
#define MNAD_XO 32   // column offset of x = 0 (row starts 128-byte aligned)
#define MNAD_ZPAD 8  // extra zero rows below the grid (rows a partial tile may read)

namespace mnad {

// Source-side d2u storage of a shot.  F32 / F16 / BF16: K snapshots kept
// (forward writes, adjoint reads); SLOT: one FP32 snapshot handed from a
// reverse-reconstruction shot to its adjoint in the same launch sequence.
enum SnapFmt : int { SNAP_F32 = 0, SNAP_F16 = 1, SNAP_BF16 = 2, SNAP_SLOT = 3 };
// FWD: forward modelling; ADJ: receiver back-propagation with imaging; REC:
// reverse reconstruction of a forward's source wavefield inside the CPML
// (from its last two levels and its saved boundary ring, mnad_ring.h).
enum Mode : int { MODE_FWD = 0, MODE_ADJ = 1, MODE_REC = 2 };

// Everything one shot needs inside the step kernel.  Lives in device memory;
// one entry per shot of a batch.
struct ShotDev {
    // Everything a step kernel needs before its first field load sits in the
    // first 192 bytes (three scalar-cache lines), so a wave fetches it in one
    // round of scalar loads.
    int nz, nx, pitch, npml;        // padded grid
    int nt;                         // number of steps
    int mode;                       // MODE_FWD / MODE_ADJ
    int t0;                         // global step index of this shot's step 0
    int ntx;                        // tiles along x (for tile -> (tj, ti))
    // four-column interior tiles (step kernel 6): one-point tiles are numbered
    // below nt1, interior tiles from nt1 on, nx4 per band of rows; they update
    // columns [npml, x4hi) of rows [npml, nz - npml), starting at x4lo
    int nt1, nx4, x4lo, x4hi;
    int nlev;
    int nrec;                       // receivers (unique points)
    int src_j, src_i;               // source (forward), -1 if none
    const float* rx;                // a / dx^2 per point, grid storage
    float* f[4][3];                 // state [level][U, Vs, Ws], grid storage: time level
                                    // t lives in f[t % nlev]; nlev = 2, or 4 for the
                                    // two-steps-per-launch kernel
    const int* rrow;                // [nz]: offset into rcol, or -1
    const int* rcol;                // [nrows_with_receivers * nx]: u or -1
    float kap;                      // rz / rx = dx^2 / dz^2
    float inv_dt2;                  // 1 / dt^2 (snapshot d2u)
    IncrC ic;                       // increment weights for this kap (mnad_incr)
    // CPML, strip-compact: x strips = 2*npml columns x nz rows,
    // z strips = 2*npml rows of `pitch` floats (columns 0 .. nx-1 used).
    float* psx;                     // [9][nz * 2 npml]   (field*3 + level)
    float* psz;                     // [9][2 npml * pitch]
    const float* cfx;               // [3][nz * 2 npml]   c1, h*c2 (signed), c3
    const float* cfz;               // [3][2 npml * pitch]
    const float* src_amp;           // [nt + 1]: a_src * w(s + 1); [nt] = 0
    float* gather;                  // forward: [nt][nrec]
    const float* radj;              // adjoint: [nt][nrec] a_u * d(tt + 1)
    // snapshots (forward) / imaging (adjoint)
    int T, K, shift;                // T = 0: off; shift = ncol - K
    int sfmt;                       // SnapFmt
    float sscale, sinv;             // f16 scale and its inverse
    void* snaps;                    // [K][nzi * nxi]
    float* illum;                   // [nzi * nxi]
    float* img;                     // [nzi * nxi] Kahan sum
    float* imgc;                    // [nzi * nxi] Kahan compensation
    // Saved boundary ring (mnad_ring.h).  FWD with rord > 0 codes the new
    // level of every ring cell; REC decodes it back, one level per step.
    // REC step s works on the forward's levels: current ntf - 1 - s, the
    // next one (read as "previous"), and writes level ntf - 2 - s.
    int rord;                       // predictor order P (2..4); 0: no ring
    int nring;                      // ring cells per field
    int rfmt;                       // code format: SNAP_F16 or SNAP_BF16
    int ntf;                        // REC: the forward's number of steps
    unsigned short* rq;             // [ntf][3][nring] codes; level L at row L - 1
    long long* rst;                 // [P][3][nring] decoded values, level L in slot L % P
    double rsc, rsi;                // fixed-point unit 2^-E: rsc = 2^E, rsi = 2^-E
    double rcs, rci;                // code scale (a power of two) and its inverse
    int kap1;                       // dx == dz (kap exactly 1): ic equals incr_consts_of(1.0)
    int rlo, rhi;                   // first and last row holding a receiver (rlo > rhi: none)
    unsigned pbytes;                // field plane stride in bytes: plane 3 level + field at
                                    // f[0][0] + that * pbytes (one allocation); 0 if not so
};

// Bytes from one field plane to the next in a shot's state allocation: the
// plane rounded up to 256 bytes.
MNAD_HD inline size_t plane_stride_bytes(int nz, int pitch)
{
    return ((size_t)(nz + 2 * MNAD_ZO + MNAD_ZPAD) * (size_t)pitch * 4 + 255) & ~(size_t)255;
}

static_assert(offsetof(ShotDev, rcol) + sizeof(void*) <= 192, "hot descriptor fields");

// Step kernel 13: a tile of `w` columns from x0 and `seg` rows from zA runs the
// interior path when every load of its march lies in the grid (four cells of
// halo) and every level-s+1 halo point (two cells) lies outside the CPML.
MNAD_HD inline bool q_interior_tile(int nx, int nz, int npml, int x0, int zA, int w, int seg)
{
    const int bw = npml + 2 > 4 ? npml + 2 : 4;
    return x0 >= bw && x0 + w <= nx - bw && zA >= bw && zA + seg <= nz - bw;
}

struct TileRef {
    int shot;   // -1: padding entry, the block exits
    int tile;   // tile index within the shot: tj * ntx + ti
};

// One block of the resident step kernel (kernel 9): rows r0 .. r0 + bh - 1 and
// columns c0 .. c0 + 63 of one shot, owned by one workgroup for a whole launch.
struct ResBlock {
    int shot, r0, c0, pad;
    int nbr[8];   // blocks holding the eight surrounding cells, -1 past the grid
};

}  // namespace mnad
