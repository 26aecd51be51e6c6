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

// Increments of (U, Vs, Ws) at one point (without 2F - F_prev and without
// the CPML term).  u, v, w: neighbourhoods of U, Vs, Ws; vww/vee: Vs at
// (j, i-2)/(j, i+2); wnn/wss: Ws at (j-2, i)/(j+2, i).
MNAD_HD void mnad_incr(float rx, float rz, const Nbh& u, const Nbh& v, float vww,
                       float vee, const Nbh& w, float wnn, float wss,
                       float& du, float& dv, float& dw)
{
    const float rx2 = rx * rx, rz2 = rz * rz, rxz = rx * rz;

    // ---- U: Auu_gen.m, Auv_gen.m, Auw_gen.m
    const float dxU = (u.e + u.w) - 2.0f * u.c;
    const float dzU = (u.n + u.s) - 2.0f * u.c;
    const float sE = (u.ne + u.se) - 2.0f * u.e;    // d2z U at column i+1
    const float sW = (u.nw + u.sw) - 2.0f * u.w;    // d2z U at column i-1
    const float sN = (u.ne + u.nw) - 2.0f * u.n;    // d2x U at row j-1
    const float sS = (u.se + u.sw) - 2.0f * u.s;    // d2x U at row j+1
    const float dxzU = (sE + sW) - 2.0f * dzU;

    const float D1v = v.e - v.w;
    const float D2v = vee - vww;
    const float Dcv = (v.ne + v.se) - (v.nw + v.sw);
    const float D1w = w.s - w.n;
    const float D2w = wss - wnn;
    const float Dcw = (w.se + w.sw) - (w.ne + w.nw);

    du = (24.0f / 11.0f) * (rx * dxU + rz * dzU)
       - (15.0f / 11.0f) * (rx2 * dxU + rz2 * dzU)
       + 0.5f * rxz * dxzU
       + ((20.0f / 33.0f) * rx - (47.0f / 66.0f) * rx2 - (1.0f / 6.0f) * rxz) * D1v
       + ((20.0f / 33.0f) * rz - (47.0f / 66.0f) * rz2 - (1.0f / 6.0f) * rxz) * D1w
       + (1.0f / 12.0f) * rxz * (Dcv + Dcw)
       - ((1.0f / 132.0f) * rx - (1.0f / 66.0f) * rx2) * D2v
       - ((1.0f / 132.0f) * rz - (1.0f / 66.0f) * rz2) * D2w;

    // ---- Vs: Avv_gen.m, Avu_gen.m, Avw_gen.m (times dx)
    const float D1u = u.e - u.w;
    const float dzD1u = sE - sW;
    const float dxV = (v.e + v.w) - 2.0f * v.c;
    const float dzV = (v.n + v.s) - 2.0f * v.c;
    const float dxzV = ((v.ne + v.se) - 2.0f * v.e) + ((v.nw + v.sw) - 2.0f * v.w) - 2.0f * dzV;
    const float CxW = (w.se + w.nw) - (w.ne + w.sw);
    dv = -15.0f * (rx - rx2) * (v.c + 0.5f * D1u)
       + (-0.5f * rz + 0.5f * rz2 - 1.25f * rxz) * dzD1u
       + (rz - 2.5f * rxz) * dzV
       - (1.5f * rx - 2.5f * rx2) * dxV
       - 0.25f * rxz * dxzV
       + 0.25f * (rz2 - rz) * CxW;

    // ---- Ws: Aww_gen.m, Awu_gen.m, Awv_gen.m (times dz)
    const float D1uz = u.s - u.n;
    const float dxD1uz = sS - sN;
    const float dxW = (w.e + w.w) - 2.0f * w.c;
    const float dzW = (w.n + w.s) - 2.0f * w.c;
    const float dxzW = ((w.se + w.sw) - 2.0f * w.s) + ((w.ne + w.nw) - 2.0f * w.n) - 2.0f * dxW;
    const float CxV = (v.se + v.nw) - (v.ne + v.sw);
    dw = -15.0f * (rz - rz2) * (w.c + 0.5f * D1uz)
       + (-0.5f * rx + 0.5f * rx2 - 1.25f * rxz) * dxD1uz
       + (rx - 2.5f * rxz) * dxW
       - (1.5f * rz - 2.5f * rz2) * dzW
       - 0.25f * rxz * dxzW
       + 0.25f * (rx2 - rx) * CxV;
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
