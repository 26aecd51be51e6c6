// mnad_stencil.h -- per-point arithmetic of the MNAD coupled update in the
// dimensionless, difference form used by the gfx950 kernels.
//
// The reference applies nine sparse operators (SRC/MNAD/*_gen.m) whose
// entries are polynomials a*alpha + b*beta of a = v^2 dt^2 and b = a^2 at the
// OUTPUT point, with constants depending only on dx, dz.  Grouping the
// entries by those constants turns every product A*F into sums of undivided
// differences of F, times rx = a/dx^2, rz = a/dz^2 and their products.  The
// gradient fields are carried scaled, Vs = dx*V and Ws = dz*W, so every
// coefficient is a pure number.  The update is then
//     F_new = 2 F + incr(F) - F_prev,
// where incr never contains the O(1) centre coefficient "2 - O(a)" of
// Auu_gen.m:44, Avv_gen.m:43, Aww_gen.m:45; in FP32 that keeps the rounding of
// the operator at the level of the O(a) terms instead of the O(1) term.
// Exactly equal, in exact arithmetic, to fwd_PDE_exact_MNAD.m:114-116.
//
// Neighbour naming: N = (j-1, i), S = (j+1, i), W = (j, i-1), E = (j, i+1),
// NE = (j-1, i+1) ... ; j is the z (row) index, i the x (column) index.
#pragma once

#ifdef __HIPCC__
#define MNAD_HD __host__ __device__ __forceinline__
#else
#define MNAD_HD inline
#endif

namespace mnad {

// 3x3 neighbourhood plus the two +-2 points a field needs.
struct Nbh {
    float c, n, s, e, w, ne, nw, se, sw;
};

// Every operator entry is alpha*rx + beta*rx^2 with constants that depend only
// on kap = rz/rx = dx^2/dz^2, so each increment is  rx * (A + rx * B)  with A
// and B fixed linear combinations of undivided differences.  IncrC holds
// those combination weights for one grid (filled once per shot, FP64 ->
// FP32).  Derivation from the entries of the *_gen.m operators, with
// k = kap:
//   U: (24/11)(rx dxU + rz dzU) - (15/11)(rx^2 dxU + rz^2 dzU) + rx rz dxzU/2
//      + ((20/33) rx - (47/66) rx^2 - rx rz/6) D1v + (same with z) D1w
//      + (rx rz/12)(Dcv + Dcw) - (rx/132 - rx^2/66) D2v - (same with z) D2w
//   Vs: -15 (rx - rx^2)(Vs + D1u/2) + (-rz/2 + rz^2/2 - 5 rx rz/4) dz D1u
//      + (rz - 5 rx rz/2) dzV - (3 rx/2 - 5 rx^2/2) dxV - (rx rz/4) dxzV
//      + (rz^2 - rz)/4 CxW
//   Ws: the same with x and z exchanged (CxV for CxW).
struct IncrC {
    float au[6], bu[8];   // U:  dxU, dzU, D1v, D1w, D2v, D2w | dxU, dzU, dxzU, D1v, D1w, Dcv+Dcw, D2v, D2w
    float av[5], bv[6];   // Vs: X, dzD1u, dzV, dxV, CxW      | X, dzD1u, dzV, dxV, dxzV, CxW
    float aw[5], bw[6];   // Ws: Y, dxD1uz, dxW, dzW, CxV     | Y, dxD1uz, dxW, dzW, dxzW, CxV
};

// The weights for kap = k (dx^2 / dz^2), each formed in double and rounded
// once to float.  constexpr: incr_consts_k1() below is the same evaluation at
// compile time, for step kernels that take the weights as literals when
// dx == dz.
MNAD_HD constexpr IncrC incr_consts_of(double k)
{
    const double k2 = k * k;
    const double au[6] = {24.0 / 11, 24.0 / 11 * k, 20.0 / 33, 20.0 / 33 * k, -1.0 / 132, -k / 132};
    const double bu[8] = {-15.0 / 11, -15.0 / 11 * k2, k / 2, -(47.0 / 66 + k / 6),
                          -(47.0 / 66 * k2 + k / 6), k / 12, 1.0 / 66, k2 / 66};
    const double av[5] = {-15.0, -0.5 * k, k, -1.5, -0.25 * k};
    const double bv[6] = {15.0, 0.5 * k2 - 1.25 * k, -2.5 * k, 2.5, -0.25 * k, 0.25 * k2};
    const double aw[5] = {-15.0 * k, -0.5, 1.0, -1.5 * k, -0.25};
    const double bw[6] = {15.0 * k2, 0.5 - 1.25 * k, -2.5 * k, 2.5 * k2, -0.25 * k, 0.25};
    IncrC c{};
    for (int i = 0; i < 6; ++i) c.au[i] = (float)au[i];
    for (int i = 0; i < 8; ++i) c.bu[i] = (float)bu[i];
    for (int i = 0; i < 5; ++i) c.av[i] = (float)av[i];
    for (int i = 0; i < 6; ++i) c.bv[i] = (float)bv[i];
    for (int i = 0; i < 5; ++i) c.aw[i] = (float)aw[i];
    for (int i = 0; i < 6; ++i) c.bw[i] = (float)bw[i];
    return c;
}

inline void incr_consts(double k, IncrC& c) { c = incr_consts_of(k); }

// Increments of (U, Vs, Ws) at one point (without 2F - F_prev and without
// the CPML term).  u, v, w: neighbourhoods of U, Vs, Ws; vww/vee: Vs at
// (j, i-2)/(j, i+2); wnn/wss: Ws at (j-2, i)/(j+2, i).
MNAD_HD void mnad_incr(const IncrC& c, float rx, const Nbh& u, const Nbh& v, float vww,
                       float vee, const Nbh& w, float wnn, float wss,
                       float& du, float& dv, float& dw)
{
    const float dxU = (u.e + u.w) - 2.0f * u.c;
    const float dzU = (u.n + u.s) - 2.0f * u.c;
    const float sE = (u.ne + u.se) - 2.0f * u.e;    // d2z U at column i+1
    const float sW = (u.nw + u.sw) - 2.0f * u.w;    // d2z U at column i-1
    const float sN = (u.ne + u.nw) - 2.0f * u.n;    // d2x U at row j-1
    const float sS = (u.se + u.sw) - 2.0f * u.s;    // d2x U at row j+1
    const float dxzU = (sE + sW) - 2.0f * dzU;
    const float D1v = v.e - v.w, D2v = vee - vww, Dcv = (v.ne + v.se) - (v.nw + v.sw);
    const float D1w = w.s - w.n, D2w = wss - wnn, Dcw = (w.se + w.sw) - (w.ne + w.nw);
    const float Au = c.au[0] * dxU + c.au[1] * dzU + c.au[2] * D1v + c.au[3] * D1w +
                     c.au[4] * D2v + c.au[5] * D2w;
    const float Bu = c.bu[0] * dxU + c.bu[1] * dzU + c.bu[2] * dxzU + c.bu[3] * D1v +
                     c.bu[4] * D1w + c.bu[5] * (Dcv + Dcw) + c.bu[6] * D2v + c.bu[7] * D2w;
    du = rx * (Au + rx * Bu);

    const float X = v.c + 0.5f * (u.e - u.w);
    const float dzD1u = sE - sW;
    const float dxV = (v.e + v.w) - 2.0f * v.c;
    const float dzV = (v.n + v.s) - 2.0f * v.c;
    const float dxzV = ((v.ne + v.se) - 2.0f * v.e) + ((v.nw + v.sw) - 2.0f * v.w) - 2.0f * dzV;
    const float CxW = (w.se + w.nw) - (w.ne + w.sw);
    const float Av = c.av[0] * X + c.av[1] * dzD1u + c.av[2] * dzV + c.av[3] * dxV + c.av[4] * CxW;
    const float Bv = c.bv[0] * X + c.bv[1] * dzD1u + c.bv[2] * dzV + c.bv[3] * dxV +
                     c.bv[4] * dxzV + c.bv[5] * CxW;
    dv = rx * (Av + rx * Bv);

    const float Y = w.c + 0.5f * (u.s - u.n);
    const float dxD1uz = sS - sN;
    const float dxW = (w.e + w.w) - 2.0f * w.c;
    const float dzW = (w.n + w.s) - 2.0f * w.c;
    const float dxzW = ((w.se + w.sw) - 2.0f * w.s) + ((w.ne + w.nw) - 2.0f * w.n) - 2.0f * dxW;
    const float CxV = (v.se + v.nw) - (v.ne + v.sw);
    const float Aw = c.aw[0] * Y + c.aw[1] * dxD1uz + c.aw[2] * dxW + c.aw[3] * dzW + c.aw[4] * CxV;
    const float Bw = c.bw[0] * Y + c.bw[1] * dxD1uz + c.bw[2] * dxW + c.bw[3] * dzW +
                     c.bw[4] * dxzW + c.bw[5] * CxV;
    dw = rx * (Aw + rx * Bw);
}

// One CPML memory triple (Ux123_gen.m:9-15, Uz123_gen.m:28-34), in the
// undivided scaling psi~ = h^2 psi: fd/f0/fa are the field at -1/0/+1 along
// the strip's axis, hc2 = h*c2 with the side's sign folded in (+ on the
// left/top side, - on the right/bottom side).  Returns the new psi1~; the
// update contributes -r*psi1~ (r = rx or rz) to the field's increment.
MNAD_HD float mnad_psi(float& p1, float& p2, float& p3, float fd, float f0,
                       float fa, float c1, float hc2, float c3)
{
    const float D = fd - fa;
    const float L = (fd - 2.0f * f0) + fa;
    const float hD = 0.5f * hc2 * D;
    p3 = c1 * p3 + hD;
    p2 = c1 * p2 + hc2 * D + c3 * (L - p3);
    p1 = c1 * p1 + hD + c3 * (2.0f * L - p2);
    return p1;
}

}  // namespace mnad
