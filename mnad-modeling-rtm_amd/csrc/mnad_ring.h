// mnad_ring.h -- the saved source-wavefield boundary of an RTM shot: layout
// of the boundary ring and its closed-loop predictive 16-bit code.  Plain
// arithmetic shared by the HIP kernels and the host tests.
//
// What it replaces.  The reference keeps full-interior d2u snapshots of the
// source wavefield every T steps (fwd_PDE_exact_MNAD.m:126-143, on disk as
// ./Wavefields/UU_<ss>.mat, read back at Born_adj_MNAD.m:33,127-134).  Here
// the forward saves, every step, only (U, Vs, Ws) on a ring of two cells
// around the reconstruction region R = the interior without the CPML strips
// (rows [npml, nz - npml), columns [npml, nx - npml)).  Inside R the CPML
// memories are zero, so the update F(t+1) = 2 F(t) + incr(F(t)) - F(t-1)
// runs backwards exactly in exact arithmetic, F(t-1) = 2 F(t) + incr(F(t)) -
// F(t+1); the stencil reaches two cells, so the ring supplies every
// neighbour outside R.  The reverse pass starts from the forward's last two
// levels and recomputes d2u = incr/dt^2 at the snapshot steps.
//
// Ring layout (cell index r, per field):
//   top band     rows npml-2, npml-1,      columns [npml-2, nx-npml+2): r in [0, 2 wb)
//   bottom band  rows nz-npml, nz-npml+1,  the same columns:             r in [2 wb, 4 wb)
//   side bands   rows [npml, nz-npml), columns npml-2, npml-1 (q = 0, 1)
//                and nx-npml, nx-npml+1 (q = 2, 3):  r = 4 wb + 4 (j - npml) + q
// with wb = nx - 2 npml + 4; nring = 4 (wb + nz - 2 npml).  npml >= 2.
//
// Code.  A ring value is stored with 8 (BF16) or 11 (F16) significant bits,
// and the reverse pass's d2u takes a second time difference of what it is
// given, which would amplify that rounding about 4/(w dt)^2 times.  So the
// forward stores the residual of a P-th order predictor instead (closed
// loop, the predictor runs on the decoded values):
//     pred(L)  = sum_i c_i G^(L - i)          c = (2,-1), (3,-3,1), (4,-6,4,-1)
//     code(L)  = fmt16((G(L) - pred(L)) * rcs)
//     G^(L)    = pred(L) + value(code(L))
// so |G^(L) - G(L)| is the rounding of one residual of size ~(w dt)^P |G|.
// The decoded values G^ are integers in units of 2^-E (|G^| < 2^62), so the
// recurrence runs backwards bit for bit:
//     G^(L - P) = c_P (G^(L) - value(code(L)) - sum_{i<P} c_i G^(L - i)),
// and the reverse pass meets exactly the values the forward's predictor saw.
// The state is P integers per cell and field, slot L mod P holding G^(L);
// levels below 1 are 0 (the fields start at rest).
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

#ifdef __HIPCC__
#define MNAD_RHD __host__ __device__ __forceinline__
#else
#define MNAD_RHD inline
#endif

namespace mnad {

constexpr int RING_MAX_ORDER = 4;

// Ring cell of padded grid point (j, i), or -1 if (j, i) is not on the ring.
MNAD_RHD int ring_index(int j, int i, int nz, int nx, int np)
{
    const int wb = nx - 2 * np + 4;
    if (i < np - 2 || i >= nx - np + 2 || j < np - 2 || j >= nz - np + 2) return -1;
    if (j < np) return (j - (np - 2)) * wb + (i - (np - 2));
    if (j >= nz - np) return 2 * wb + (j - (nz - np)) * wb + (i - (np - 2));
    if (i < np) return 4 * wb + 4 * (j - np) + (i - (np - 2));
    if (i >= nx - np) return 4 * wb + 4 * (j - np) + 2 + (i - (nx - np));
    return -1;
}

MNAD_RHD int ring_cells(int nz, int nx, int np) { return 4 * ((nx - 2 * np + 4) + (nz - 2 * np)); }

// The grid point (j, i) of ring cell r (the inverse of ring_index).
MNAD_RHD void ring_point(int r, int nz, int nx, int np, int* j, int* i)
{
    const int wb = nx - 2 * np + 4;
    if (r < 2 * wb) {
        *j = np - 2 + r / wb;
        *i = np - 2 + r % wb;
    } else if (r < 4 * wb) {
        const int q = r - 2 * wb;
        *j = nz - np + q / wb;
        *i = np - 2 + q % wb;
    } else {
        const int q = r - 4 * wb, c = q % 4;
        *j = np + q / 4;
        *i = c < 2 ? np - 2 + c : nx - np + (c - 2);
    }
}

// Point (j, i) inside the reconstruction region.
MNAD_RHD bool in_rec_region(int j, int i, int nz, int nx, int np)
{
    return j >= np && j < nz - np && i >= np && i < nx - np;
}

MNAD_RHD uint32_t f32_bits(float x)
{
    uint32_t b;
    memcpy(&b, &x, 4);
    return b;
}

MNAD_RHD float f32_from(uint32_t b)
{
    float x;
    memcpy(&x, &b, 4);
    return x;
}

// bfloat16, round to nearest even (NaN -> quiet NaN).
MNAD_RHD uint16_t bf16_enc(float x)
{
    uint32_t b = f32_bits(x);
    if ((b & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    b += 0x7fffu + ((b >> 16) & 1u);
    return (uint16_t)(b >> 16);
}

MNAD_RHD float bf16_dec(uint16_t h) { return f32_from((uint32_t)h << 16); }

// IEEE half, round to nearest even, subnormals kept, overflow to infinity.
MNAD_RHD uint16_t f16_enc(float x)
{
    const uint32_t b = f32_bits(x);
    const uint16_t sign = (uint16_t)((b >> 16) & 0x8000u);
    const float a = fabsf(x);
    if (a != a) return (uint16_t)(sign | 0x7e00u);
    if (!(a < 65520.0f)) return (uint16_t)(sign | 0x7c00u);
    if (a < 6.103515625e-05f)   // half subnormal: m * 2^-24 (a * 2^24 is exact)
        return (uint16_t)(sign | (uint32_t)rintf(a * 16777216.0f));
    uint32_t m = f32_bits(a);
    m += 0xfffu + ((m >> 13) & 1u);
    return (uint16_t)(sign | ((((m >> 23) - 112u) << 10) | ((m >> 13) & 0x3ffu)));
}

MNAD_RHD float f16_dec(uint16_t h)
{
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    if (e == 0) {
        const float v = (float)m * 5.9604644775390625e-08f;   // 2^-24
        return sign ? -v : v;
    }
    if (e == 31) return f32_from(sign | 0x7f800000u | (m << 13));
    return f32_from(sign | ((e + 112u) << 23) | (m << 13));
}

// Predictor coefficients c_1 .. c_P (binomial, alternating).
MNAD_RHD int64_t ring_pred(int P, int64_t g1, int64_t g2, int64_t g3, int64_t g4)
{
    if (P == 2) return 2 * g1 - g2;
    if (P == 3) return 3 * (g1 - g2) + g3;
    return 4 * (g1 + g3) - 6 * g2 - g4;
}

// Integer value (units 2^-E) of a stored code; `fmt16` 1 = F16, 2 = BF16;
// rci = 1 / rcs, a power of two.
MNAD_RHD int64_t ring_value(uint16_t code, int fmt16, double rci)
{
    const double x = (double)(fmt16 == 1 ? f16_dec(code) : bf16_dec(code)) * rci;
    const double r = rint(x);
    return fabs(r) < 4.0e18 ? (int64_t)r : 0;   // non-finite input (an unstable shot)
}

// Forward: code of value x at level L given the decoded G^(L-1) .. G^(L-P)
// (g[0] .. g[P-1], zero beyond P); returns the code and sets *gL = G^(L).
// rsc = 2^E (fixed-point scale), rcs = residual scale of the code.
MNAD_RHD uint16_t ring_encode(int P, const int64_t* g, float x, double rsc, double rcs, double rci,
                              int fmt16, int64_t* gL)
{
    const int64_t pred = ring_pred(P, g[0], g[1], P > 2 ? g[2] : 0, P > 3 ? g[3] : 0);
    const double r = (double)x * rsc - (double)pred;
    const float rf = (float)(r * rcs);
    const uint16_t code = fmt16 == 1 ? f16_enc(rf) : bf16_enc(rf);
    *gL = pred + ring_value(code, fmt16, rci);
    return code;
}

// Reverse: G^(L - P) from G^(L) (gL), G^(L-1) .. G^(L-P+1) (g[0] .. g[P-2])
// and the code of level L.
MNAD_RHD int64_t ring_decode(int P, int64_t gL, const int64_t* g, uint16_t code, int fmt16,
                             double rci)
{
    const int64_t q = ring_value(code, fmt16, rci);
    if (P == 2) return -(gL - q - 2 * g[0]);
    if (P == 3) return gL - q - 3 * (g[0] - g[1]);
    return -(gL - q - 4 * g[0] + 6 * g[1] - 4 * g[2]);
}

}  // namespace mnad
