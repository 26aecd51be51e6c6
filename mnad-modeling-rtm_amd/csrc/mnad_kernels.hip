// mnad_kernels.hip -- gfx950 kernels of the MNAD hot path.
//
//   k_setup_model  FP64 set-up: a/dx^2 per point and the CPML coefficients of
//                  C_cof.m, written as FP32 strip-compact arrays.
//   k_step         one time step of one or many shots (forward or adjoint):
//                  the coupled (U, Vs, Ws) update of fwd_PDE_exact_MNAD.m:114-116
//                  with the CPML memories of Ux123_gen.m/Uz123_gen.m folded in on
//                  the strips, source injection, receiver sampling
//                  (fwd_PDE_exact_MNAD.m:111,119-123), receiver injection
//                  (Born_adj_MNAD.m:109-111), d2u snapshots + illumination
//                  (fwd_PDE_exact_MNAD.m:126-143) and the zero-lag imaging
//                  condition with compensated FP32 sums (Born_adj_MNAD.m:127-134).
//   k_advance      bumps the device-side step counter (end of a graph chunk).
//   k_make_radj    a_u * (d_obs - d_syn) per adjoint step (main_Acoustic_RTM.m:225
//                  and Born_adj_MNAD.m:110), from two device gathers.
//   k_stack        image / illumination stack of main_Acoustic_RTM.m:237-262.
//
// Injection bookkeeping.  The reference injects the source into the CURRENT
// U before the update (fwd_PDE_exact_MNAD.m:111) and the receiver data into
// the current U after it (Born_adj_MNAD.m:109-111).  Both are folded into the
// output of the step before, so no extra launch and no read/write race on U:
//   forward: step s writes  U_new(src) += a_src w(s+2)  (the next step's
//            injection); the CPML memories, which see the pre-injection U in the
//            reference, subtract it back when they read U at the source point.
//   adjoint: the injected value only survives as the next step's -U_prev, so
//            step s writes U_new(rec) -= a_rec d(tt+1).
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include <utility>

#include "mnad_device.h"
#include "mnad_stencil.h"

namespace mnad {

constexpr int BX = 64;   // tile width  (x), one wavefront per tile row
constexpr int BZ = 4;    // tile height (z)

// ------------------------------------------------------------------ set-up
// v: column-major [nz x nx] FP64 as MATLAB passes it (v(j,i) = v[j + i*nz]).
// a = v^2 dt^2 (fwd_PDE_exact_MNAD.m:39; *_gen.m use v(j,i).^2*dt^2).
// CPML (C_cof.m:4-65): sigma = 3 ln(1/R)/(2 L^3) v x^2, sigma' = 2 (...) v x,
// den = 1 + sigma dt/2, c1 = (1 - sigma dt/2)/den, c2 = sigma' dt/den,
// c3 = sigma dt/den; x = distance into the layer (h ... L), L = npml h.
__global__ void k_setup_model(const double* __restrict__ v, int nz, int nx,
                              int pitch, int npml, double dx, double dz,
                              double dt, double R, float* __restrict__ rx,
                              float* __restrict__ cfx, float* __restrict__ cfz)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= nx || j >= nz) return;
    const double vv = v[(size_t)j + (size_t)i * nz];
    const double a = vv * vv * (dt * dt);
    rx[(size_t)(j + MNAD_ZO) * pitch + (i + MNAD_XO)] = (float)(a / (dx * dx));
    if (npml <= 0) return;
    const int np2 = 2 * npml;
    if (i < npml || i >= nx - npml) {
        const double L = npml * dx;
        const double k = (3.0 * log(1.0 / R)) / (2.0 * L * L * L);
        const bool left = i < npml;
        const double x = left ? (double)(npml - i) * dx : (double)(i - (nx - npml) + 1) * dx;
        const double sig0 = k * vv;
        const double sig = sig0 * (x * x);
        const double dsig = 2.0 * sig0 * x;
        const double den = 1.0 + sig * (dt / 2.0);
        const int ix = left ? i : i - (nx - np2);
        const size_t q = (size_t)j * np2 + ix, ns = (size_t)nz * np2;
        cfx[q] = (float)((1.0 - sig * (dt / 2.0)) / den);
        cfx[ns + q] = (float)((left ? dx : -dx) * ((dsig * dt) / den));
        cfx[2 * ns + q] = (float)((sig * dt) / den);
    }
    if (j < npml || j >= nz - npml) {
        const double L = npml * dz;
        const double k = (3.0 * log(1.0 / R)) / (2.0 * L * L * L);
        const bool top = j < npml;
        const double x = top ? (double)(npml - j) * dz : (double)(j - (nz - npml) + 1) * dz;
        const double sig0 = k * vv;
        const double sig = sig0 * (x * x);
        const double dsig = 2.0 * sig0 * x;
        const double den = 1.0 + sig * (dt / 2.0);
        const int jz = top ? j : j - (nz - np2);
        const size_t q = (size_t)jz * pitch + i, ns = (size_t)np2 * pitch;
        cfz[q] = (float)((1.0 - sig * (dt / 2.0)) / den);
        cfz[ns + q] = (float)((top ? dz : -dz) * ((dsig * dt) / den));
        cfz[2 * ns + q] = (float)((sig * dt) / den);
    }
}

// ------------------------------------------------------------ step kernel
// Snapshot element store / load: a two-way branch on the storage width,
// the half / bfloat16 choice as a select (see snap_image_w).
__device__ __forceinline__ void store_snap(const ShotDev& d, long idx, float x)
{
    if (d.sfmt == SNAP_F32) {
        static_cast<float*>(d.snaps)[idx] = x;
    } else {
        const unsigned short h16 = __half_as_ushort(__float2half(x * d.sscale));
        const unsigned short b16 = __bfloat16_as_ushort(__float2bfloat16(x));
        static_cast<unsigned short*>(d.snaps)[idx] = d.sfmt == SNAP_F16 ? h16 : b16;
    }
}

__device__ __forceinline__ float load_snap(const ShotDev& d, long idx)
{
    if (d.sfmt == SNAP_F32) return static_cast<const float*>(d.snaps)[idx];
    const unsigned short h = static_cast<const unsigned short*>(d.snaps)[idx];
    return d.sfmt == SNAP_F16 ? __half2float(__ushort_as_half(h)) * d.sinv
                              : __bfloat162float(__ushort_as_bfloat16(h));
}

// Three CPML memory triples (U, Vs, Ws) of one strip point along one axis.
// ps: 9 planes of `ns` floats (field*3 + level), cf: 3 planes (c1, h c2, c3).
__device__ __forceinline__ void psi3(float* __restrict__ ps, const float* __restrict__ cf,
                                     long ns, long q, float r,
                                     float ud, float uc, float ua,
                                     float vd, float vc, float va,
                                     float wd, float wc, float wa,
                                     float& du, float& dv, float& dw)
{
    const float c1 = cf[q], hc2 = cf[ns + q], c3 = cf[2 * ns + q];
    float p1, p2, p3;
    p1 = ps[q]; p2 = ps[ns + q]; p3 = ps[2 * ns + q];
    du -= r * mnad_psi(p1, p2, p3, ud, uc, ua, c1, hc2, c3);
    ps[q] = p1; ps[ns + q] = p2; ps[2 * ns + q] = p3;
    p1 = ps[3 * ns + q]; p2 = ps[4 * ns + q]; p3 = ps[5 * ns + q];
    dv -= r * mnad_psi(p1, p2, p3, vd, vc, va, c1, hc2, c3);
    ps[3 * ns + q] = p1; ps[4 * ns + q] = p2; ps[5 * ns + q] = p3;
    p1 = ps[6 * ns + q]; p2 = ps[7 * ns + q]; p3 = ps[8 * ns + q];
    dw -= r * mnad_psi(p1, p2, p3, wd, wc, wa, c1, hc2, c3);
    ps[6 * ns + q] = p1; ps[7 * ns + q] = p2; ps[8 * ns + q] = p3;
}

__device__ __forceinline__ void snap_or_image(const ShotDev& d, int s, int j, int i, float du,
                                              float nu);

// Everything after the increments for one point: CPML memories on the
// strips, the new level, injections, gathers, snapshots and imaging.
__device__ __forceinline__ void finish_point(const ShotDev& d, int s, int j, int i, long p,
                                             const Nbh& u, const Nbh& v, const Nbh& w,
                                             float du, float dv, float dw, float rx, float rz,
                                             float* __restrict__ PU, float* __restrict__ PV,
                                             float* __restrict__ PW)
{
    const float pu = PU[p], pv = PV[p], pw = PW[p];
    const int np = d.npml;
    const bool inx = (i < np) || (i >= d.nx - np);
    const bool inz = (j < np) || (j >= d.nz - np);
    if (inx || inz) {
        // The CPML memories see the pre-injection U (fwd_PDE_exact_MNAD.m:95-111).
        float uc = u.c, un = u.n, us = u.s, ue = u.e, uw = u.w;
        if (d.mode == MODE_FWD && d.src_j >= 0) {
            const float a0 = d.src_amp[s];
            if (j == d.src_j) {
                if (i == d.src_i) uc -= a0;
                if (i - 1 == d.src_i) uw -= a0;
                if (i + 1 == d.src_i) ue -= a0;
            }
            if (i == d.src_i) {
                if (j - 1 == d.src_j) un -= a0;
                if (j + 1 == d.src_j) us -= a0;
            }
        }
        if (inx) {
            const long ns = (long)d.nz * 2 * np;
            const int ix = (i < np) ? i : i - (d.nx - 2 * np);
            const long q = (long)j * 2 * np + ix;
            psi3(d.psx, d.cfx, ns, q, rx, uw, uc, ue, v.w, v.c, v.e, w.w, w.c, w.e, du, dv, dw);
        }
        if (inz) {
            const long ns = (long)2 * np * d.pitch;
            const int jz = (j < np) ? j : j - (d.nz - 2 * np);
            const long q = (long)jz * d.pitch + i;
            psi3(d.psz, d.cfz, ns, q, rz, un, uc, us, v.n, v.c, v.s, w.n, w.c, w.s, du, dv, dw);
        }
    }

    float nu = (2.0f * u.c + du) - pu;
    const float nv = (2.0f * v.c + dv) - pv;
    const float nw = (2.0f * w.c + dw) - pw;

    int r = -1;
    if (d.nrec > 0) {
        const int off = d.rrow[j];
        if (off >= 0) r = d.rcol[off + i];
    }
    if (d.mode == MODE_FWD) {
        if (j == d.src_j && i == d.src_i) nu += d.src_amp[s + 1];
        if (r >= 0) d.gather[(long)s * d.nrec + r] = u.c;
    } else if (r >= 0) {
        nu -= d.radj[(long)s * d.nrec + r];
    }
    PU[p] = nu;
    PV[p] = nv;
    PW[p] = nw;
    if (d.T > 0 && !inx && !inz) snap_or_image(d, s, j, i, du, nu);
}

// d2u snapshot + illumination (forward) or imaging (adjoint) at an interior
// point on the steps where the schedule takes them.
__device__ __forceinline__ void snap_or_image(const ShotDev& d, int s, int j, int i, float du,
                                              float nu)
{
    const int np = d.npml;
    {
        const int nxi = d.nx - 2 * np;
        const long q = (long)(j - np) * nxi + (i - np);
        const long ni = (long)(d.nz - 2 * np) * nxi;
        if (d.mode == MODE_FWD) {
            if ((s + 1) % d.T == 0) {
                // d2u = (nU - 2U + pU)/dt^2 = du/dt^2; snapshot k = (s+1)/T - 1.
                const int k = (s + 1) / d.T - 1;
                const float d2u = du * d.inv_dt2;
                d.illum[q] += d2u * d2u;
                store_snap(d, (long)k * ni + q, d2u);
            }
        } else {
            const int tt = d.nt - s;
            if (tt % d.T == 0) {
                // UU(:, ncol - ttt + 1) = snapshot column shift + tt/T (1-based).
                const int col = d.shift + tt / d.T;
                if (col <= d.K) {
                    const float x = load_snap(d, (long)(col - 1) * ni + q) * nu;
                    const float sum = d.img[q], c = d.imgc[q];
                    const float y = x - c;
                    const float t = sum + y;
                    d.imgc[q] = (t - sum) - y;
                    d.img[q] = t;
                }
            }
        }
    }
}

// One step for every tile in `tiles`.  Block = BX x BZR threads; each thread
// updates RZ consecutive rows of one column, so a tile is BX x (BZR*RZ)
// points and the column values shared by those rows are loaded once.
// Local step of a shot: s = *base + rel - t0 (0-based); tiles of shots outside
// [0, nt) exit at once, so a graph chunk may run past a shot's end.
// Storage keeps MNAD_ZO + MNAD_ZPAD zero rows below the grid, so the loads of
// a partial bottom tile stay inside the allocation (RZ <= MNAD_ZPAD).
template <int RZ, int BZR = BZ>
__global__ __launch_bounds__(BX * BZR) void k_step(const ShotDev* __restrict__ shots,
                                                  const TileRef* __restrict__ tiles,
                                                  const int* __restrict__ base, int rel)
{
    static_assert(RZ >= 1 && RZ <= MNAD_ZPAD, "rows per thread");
    const TileRef tr = tiles[blockIdx.x];
    if (tr.shot < 0) return;
    const ShotDev& d = shots[tr.shot];
    const int s = *base + rel - d.t0;
    if (s < 0 || s >= d.nt) return;
    const int tj = tr.tile / d.ntx;
    const int ti = tr.tile - tj * d.ntx;
    const int i = ti * BX + threadIdx.x;
    const int j0 = (tj * BZR + threadIdx.y) * RZ;
    if (i >= d.nx || j0 >= d.nz) return;

    const int cb = s & 1;
    const float* __restrict__ U = d.f[cb][0];
    const float* __restrict__ V = d.f[cb][1];
    const float* __restrict__ W = d.f[cb][2];
    float* __restrict__ PU = d.f[cb ^ 1][0];
    float* __restrict__ PV = d.f[cb ^ 1][1];
    float* __restrict__ PW = d.f[cb ^ 1][2];
    const int P = d.pitch;
    const long p0 = (long)(j0 + MNAD_ZO) * P + (i + MNAD_XO);

    // Column windows: rows j0-1 .. j0+RZ (U, V, W sides), j0-2 .. j0+RZ+1 (W centre).
    float ua[RZ + 2][3], va[RZ + 2][3], vf[RZ][2], wc[RZ + 4], ws[RZ + 2][2];
#pragma unroll
    for (int k = 0; k < RZ + 2; ++k) {
        const long q = p0 + (long)(k - 1) * P;
        ua[k][0] = U[q - 1]; ua[k][1] = U[q]; ua[k][2] = U[q + 1];
        va[k][0] = V[q - 1]; va[k][1] = V[q]; va[k][2] = V[q + 1];
        ws[k][0] = W[q - 1]; ws[k][1] = W[q + 1];
    }
#pragma unroll
    for (int k = 0; k < RZ + 4; ++k) wc[k] = W[p0 + (long)(k - 2) * P];
#pragma unroll
    for (int k = 0; k < RZ; ++k) {
        vf[k][0] = V[p0 + (long)k * P - 2];
        vf[k][1] = V[p0 + (long)k * P + 2];
    }

#pragma unroll
    for (int r = 0; r < RZ; ++r) {
        const int j = j0 + r;
        if (j >= d.nz) break;
        const long p = p0 + (long)r * P;
        Nbh u, v, w;
        u.nw = ua[r][0];     u.n = ua[r][1];     u.ne = ua[r][2];
        u.w = ua[r + 1][0];  u.c = ua[r + 1][1]; u.e = ua[r + 1][2];
        u.sw = ua[r + 2][0]; u.s = ua[r + 2][1]; u.se = ua[r + 2][2];
        v.nw = va[r][0];     v.n = va[r][1];     v.ne = va[r][2];
        v.w = va[r + 1][0];  v.c = va[r + 1][1]; v.e = va[r + 1][2];
        v.sw = va[r + 2][0]; v.s = va[r + 2][1]; v.se = va[r + 2][2];
        w.nw = ws[r][0];     w.n = wc[r + 1];    w.ne = ws[r][1];
        w.w = ws[r + 1][0];  w.c = wc[r + 2];    w.e = ws[r + 1][1];
        w.sw = ws[r + 2][0]; w.s = wc[r + 3];    w.se = ws[r + 2][1];
        const float rx = d.rx[p];
        const float rz = rx * d.kap;
        float du, dv, dw;
        mnad_incr(d.ic, rx, u, v, vf[r][0], vf[r][1], w, wc[r], wc[r + 4], du, dv, dw);
        finish_point(d, s, j, i, p, u, v, w, du, dv, dw, rx, rz, PU, PV, PW);
    }
}

// ------------------------------------------------------ wave-row step kernel
// Same update as k_step<1, BZR>, laid out for the gfx950 addressing modes: a
// wavefront owns 64 consecutive points of ONE grid row, so every origin it
// needs (field row, CPML plane row, snapshot row) is wave-uniform and lives
// in SGPRs, and every load or store is
//     global_load_dword v, v_lane, s[origin] offset:imm
// with a single lane-offset VGPR shared by all of them (the compiler also
// merges the -1/0/+1 neighbours of a row into one dwordx3 load).  The
// descriptors hold generic pointers; they are cast to the global address
// space, which would otherwise compile to flat instructions with 64-bit
// per-lane address arithmetic.  All CPML loads of a point are issued before
// its stores.
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) int gint;
typedef __attribute__((address_space(1))) unsigned short gu16;

__device__ __forceinline__ gfloat* gp(float* p) { return (gfloat*)p; }
__device__ __forceinline__ const gfloat* gp(const float* p) { return (const gfloat*)p; }
__device__ __forceinline__ const gint* gp(const int* p) { return (const gint*)p; }

// A wave-uniform pointer made opaque (readfirstlane of both halves), so that
// the compiler keeps it in SGPRs and does not fold a per-lane index into
// 64-bit VALU address arithmetic.
template <class T>
__device__ __forceinline__ T* uni(T* p)
{
    const unsigned long long a = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
    return (T*)(((unsigned long long)hi << 32) | lo);
}

// Element at a 32-bit byte offset from a global base (SGPR base + VGPR offset).
typedef __attribute__((address_space(1))) char gchar;
__device__ __forceinline__ gfloat& at(gfloat* base, unsigned off)
{
    return *(gfloat*)((gchar*)base + off);
}
__device__ __forceinline__ const gfloat& at(const gfloat* base, unsigned off)
{
    return *(const gfloat*)((const gchar*)base + off);
}

// CPML state of one strip point along one axis: coefficients (c1, h c2, c3)
// and the nine memories (field * 3 + level).  ps/cf are the plane origins of
// the point's strip row (wave-uniform), b = 4 q with q the lane's column
// within that row (q < 2 npml or q < 64, so b fits the 32-bit offset of the
// SGPR-base form).
struct PsiW {
    float c[3], p[9];
};

__device__ __forceinline__ void psi_load(PsiW& m, const gfloat* ps, const gfloat* cf, long ns,
                                         unsigned b)
{
#pragma unroll
    for (int k = 0; k < 3; ++k) m.c[k] = at(uni(cf + k * ns), b);
#pragma unroll
    for (int k = 0; k < 9; ++k) m.p[k] = at(uni(ps + k * ns), b);
}

__device__ __forceinline__ void psi_step(PsiW& m, float r, float ud, float uc, float ua, float vd,
                                         float vc, float va, float wd, float wc, float wa,
                                         float& du, float& dv, float& dw)
{
    du -= r * mnad_psi(m.p[0], m.p[1], m.p[2], ud, uc, ua, m.c[0], m.c[1], m.c[2]);
    dv -= r * mnad_psi(m.p[3], m.p[4], m.p[5], vd, vc, va, m.c[0], m.c[1], m.c[2]);
    dw -= r * mnad_psi(m.p[6], m.p[7], m.p[8], wd, wc, wa, m.c[0], m.c[1], m.c[2]);
}

__device__ __forceinline__ void psi_store(const PsiW& m, gfloat* ps, long ns, unsigned b)
{
#pragma unroll
    for (int k = 0; k < 9; ++k) at(uni(ps + k * ns), b) = m.p[k];
}

// d2u snapshot + illumination (forward) or imaging (adjoint) of one interior
// point, on the steps where the schedule takes them.  qrow: wave-uniform
// offset of the point's interior row, q: its interior column.
__device__ __forceinline__ void snap_image_w(const ShotDev& d, int s, long qrow, unsigned q,
                                             float du, float nu)
{
    // Two-way branches on the storage width only (32 or 16 bits) with the
    // half/bfloat16 choice as a select: written as a three-way if/else chain,
    // hipcc compiled the bfloat16 store path so that the illumination update
    // after it read its column offset from a register no path had written.
    const long ni = (long)(d.nz - 2 * d.npml) * (d.nx - 2 * d.npml);
    // the one-slot hand-off is FP32 like the F32 snapshots
    const bool f32 = d.sfmt == SNAP_F32 || d.sfmt == SNAP_SLOT, f16 = d.sfmt == SNAP_F16;
    if (d.mode == MODE_REC) {
        // d2u of the forward's step ntf - s (1-based); on a snapshot step it
        // goes to the slot, which the adjoint reads one launch later
        if ((d.ntf - s) % d.T == 0) (gp(static_cast<float*>(d.snaps)) + qrow)[q] = du * d.inv_dt2;
        return;
    }
    if (d.mode == MODE_FWD) {
        if ((s + 1) % d.T == 0) {
            // d2u = (nU - 2U + pU)/dt^2 = du/dt^2; snapshot k = (s+1)/T - 1.
            const long k0 = (long)((s + 1) / d.T - 1) * ni + qrow;
            const float d2u = du * d.inv_dt2;
            gfloat* il = gp(d.illum) + qrow;
            il[q] += d2u * d2u;
            if (d.snaps) {   // a forward that saves its boundary ring keeps no snapshots
                if (f32) {
                    (gp(static_cast<float*>(d.snaps)) + k0)[q] = d2u;
                } else {
                    const unsigned short h16 = __half_as_ushort(__float2half(d2u * d.sscale));
                    const unsigned short b16 = __bfloat16_as_ushort(__float2bfloat16(d2u));
                    ((gu16*)d.snaps + k0)[q] = f16 ? h16 : b16;
                }
            }
        }
    } else {
        const int tt = d.nt - s;
        if (tt % d.T == 0) {
            // UU(:, ncol - ttt + 1) = snapshot column shift + tt/T (1-based).
            const int col = d.shift + tt / d.T;
            if (col <= d.K) {
                const long k0 = (d.sfmt == SNAP_SLOT ? 0L : (long)(col - 1) * ni) + qrow;
                float sn;
                if (f32) {
                    sn = (gp(static_cast<const float*>(d.snaps)) + k0)[q];
                } else {
                    const unsigned short h = ((const gu16*)d.snaps + k0)[q];
                    sn = f16 ? __half2float(__ushort_as_half(h)) * d.sinv
                             : __bfloat162float(__ushort_as_bfloat16(h));
                }
                const float x = sn * nu;
                gfloat* im = gp(d.img) + qrow;
                gfloat* ic = gp(d.imgc) + qrow;
                const float sum = im[q], c = ic[q];
                const float y = x - c;
                const float t = sum + y;
                ic[q] = (t - sum) - y;
                im[q] = t;
            }
        }
    }
}

// ------------------------------------------------ saved boundary ring
typedef __attribute__((address_space(1))) long long gi64;

// Forward, ring cell ri: code level L of (U, Vs, Ws) (mnad_ring.h).
template <int P>
__device__ __forceinline__ void ring_put_p(const ShotDev& d, int L, unsigned ri, float nu,
                                           float nv, float nw)
{
    const long nr = d.nring;
    gi64* st = (gi64*)d.rst;
    gu16* rq = (gu16*)d.rq;
    const float x[3] = {nu, nv, nw};
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        const long cell = f * nr + ri;
        int64_t g[RING_MAX_ORDER] = {0, 0, 0, 0};
#pragma unroll
        for (int i = 1; i <= P; ++i) g[i - 1] = st[(long)(((L - i) % P + P) % P) * 3 * nr + cell];
        int64_t gL;
        const unsigned short code = ring_encode(P, g, x[f], d.rsc, d.rcs, d.rci, d.rfmt, &gL);
        st[(long)(L % P) * 3 * nr + cell] = gL;
        rq[(long)(L - 1) * 3 * nr + cell] = code;
    }
}

__device__ __forceinline__ void ring_put(const ShotDev& d, int L, unsigned ri, float nu, float nv,
                                         float nw)
{
    if (d.rord == 2) ring_put_p<2>(d, L, ri, nu, nv, nw);
    else if (d.rord == 3) ring_put_p<3>(d, L, ri, nu, nv, nw);
    else ring_put_p<4>(d, L, ri, nu, nv, nw);
}

// Reverse, ring cell ri: with the decoded window G^(L) .. G^(L-P+1) and the
// code of level L, decode G^(L-P) into the slot of G^(L) and return level
// L - 2 (the level the reverse step writes).
template <int P>
__device__ __forceinline__ void ring_get_p(const ShotDev& d, int L, unsigned ri, float* out)
{
    const long nr = d.nring;
    gi64* st = (gi64*)d.rst;
    const gu16* rq = (const gu16*)d.rq;
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        const long cell = f * nr + ri;
        const long top = (long)(L % P) * 3 * nr + cell;
        const int64_t gL = st[top];
        int64_t g[RING_MAX_ORDER] = {0, 0, 0, 0};
#pragma unroll
        for (int i = 1; i < P; ++i) g[i - 1] = st[(long)(((L - i) % P + P) % P) * 3 * nr + cell];
        const int64_t back = ring_decode(P, gL, g, rq[(long)(L - 1) * 3 * nr + cell], d.rfmt, d.rci);
        st[top] = back;
        out[f] = (float)((double)(P == 2 ? back : g[1]) * d.rsi);
    }
}

__device__ __forceinline__ void ring_get(const ShotDev& d, int L, unsigned ri, float* out)
{
    if (d.rord == 2) ring_get_p<2>(d, L, ri, out);
    else if (d.rord == 3) ring_get_p<3>(d, L, ri, out);
    else ring_get_p<4>(d, L, ri, out);
}

// Neighbourhoods of U, Vs, Ws around (j, i0 + lane): row origins U, V, W.
__device__ __forceinline__ void load_nbh(const gfloat* U, const gfloat* V, const gfloat* W, long P,
                                         unsigned lane, Nbh& u, Nbh& v, Nbh& w, float& vww,
                                         float& vee, float& wnn, float& wss)
{
    const gfloat* a = (U - P) + lane;
    u.nw = a[-1]; u.n = a[0]; u.ne = a[1];
    a = U + lane;
    u.w = a[-1];  u.c = a[0]; u.e = a[1];
    a = (U + P) + lane;
    u.sw = a[-1]; u.s = a[0]; u.se = a[1];
    a = (V - P) + lane;
    v.nw = a[-1]; v.n = a[0]; v.ne = a[1];
    a = V + lane;
    vww = a[-2];  v.w = a[-1]; v.c = a[0]; v.e = a[1]; vee = a[2];
    a = (V + P) + lane;
    v.sw = a[-1]; v.s = a[0]; v.se = a[1];
    wnn = ((W - 2 * P) + lane)[0];
    a = (W - P) + lane;
    w.nw = a[-1]; w.n = a[0]; w.ne = a[1];
    a = W + lane;
    w.w = a[-1];  w.c = a[0]; w.e = a[1];
    a = (W + P) + lane;
    w.sw = a[-1]; w.s = a[0]; w.se = a[1];
    wss = ((W + 2 * P) + lane)[0];
}

// Reverse reconstruction (MODE_REC), one point per lane of row j: step s
// turns levels c = ntf-1-s (current) and c+1 into level c-1, written over
// c+1.  Inside the reconstruction region R the update runs backwards,
//     F(c-1) = 2 F(c) + incr(F(c)) - F(c+1)   (+ a_src w at the source,
// the forward's injection into level c+1 taken back out); R holds no CPML
// memories.  Ring cells decode level c-1 from the saved ring; every other
// point is left alone (R never reads it).  On the forward's snapshot steps
// d2u = incr/dt^2 goes to the adjoint's slot.  SKIPX4: points of the
// multi-column region belong to step_xn.
template <bool SKIPX4>
__device__ __forceinline__ void rec_one(const ShotDev& d, int s, int j, int i0, unsigned lane)
{
    const int i = i0 + (int)lane;
    if (i >= d.nx) return;
    const int np = d.npml;
    const int ri = ring_index(j, i, d.nz, d.nx, np);
    if (ri < 0) {
        if (!in_rec_region(j, i, d.nz, d.nx, np)) return;
        if (SKIPX4 && i < d.x4hi) return;
    }
    const int c = d.ntf - 1 - s;
    const bool odd = c & 1;
    const long P = d.pitch;
    const long org = (long)(j + MNAD_ZO) * P + (i0 + MNAD_XO);
    gfloat* NU = gp(odd ? d.f[0][0] : d.f[1][0]) + org;
    gfloat* NV = gp(odd ? d.f[0][1] : d.f[1][1]) + org;
    gfloat* NW = gp(odd ? d.f[0][2] : d.f[1][2]) + org;
    if (ri >= 0) {
        float o[3];
        ring_get(d, c + 1, (unsigned)ri, o);
        NU[lane] = o[0];
        NV[lane] = o[1];
        NW[lane] = o[2];
        return;
    }
    Nbh u, v, w;
    float vww, vee, wnn, wss;
    load_nbh(gp(odd ? d.f[1][0] : d.f[0][0]) + org, gp(odd ? d.f[1][1] : d.f[0][1]) + org,
             gp(odd ? d.f[1][2] : d.f[0][2]) + org, P, lane, u, v, w, vww, vee, wnn, wss);
    const float rx = (gp(d.rx) + org)[lane];
    const float pu = NU[lane], pv = NV[lane], pw = NW[lane];
    float du, dv, dw;
    mnad_incr(d.ic, rx, u, v, vww, vee, w, wnn, wss, du, dv, dw);
    float nu = (2.0f * u.c + du) - pu;
    if (j == d.src_j && i == d.src_i) nu += d.src_amp[c + 1];
    NU[lane] = nu;
    NV[lane] = (2.0f * v.c + dv) - pv;
    NW[lane] = (2.0f * w.c + dw) - pw;
    if (d.T > 0) snap_image_w(d, s, (long)(j - np) * (d.nx - 2 * np), (unsigned)(i - np), du, nu);
}

// One point per lane: the wave owns columns i0 .. i0+63 of row j.
// SKIPX4: points of the four-column region (rows npml .. nz-npml-1, columns
// npml .. x4hi-1) belong to step_xn and are left alone.
template <bool SKIPX4>
__device__ __forceinline__ void step_one(const ShotDev& d, int s, int j, int i0, unsigned lane)
{
    const int i = i0 + (int)lane;
    if (i >= d.nx) return;
    if (SKIPX4 && j >= d.npml && j < d.nz - d.npml && i >= d.npml && i < d.x4hi) return;
    const int np = d.npml;
    const bool inx = (i < np) || (i >= d.nx - np);
    const bool inz = (j < np) || (j >= d.nz - np);   // wave-uniform
    // strip-row origins: x strips [9][nz * 2 npml], z strips [9][2 npml * nx]
    const long nsx = (long)d.nz * 2 * np, rowx = (long)j * 2 * np;
    const unsigned bx = 4u * ((i < np) ? (unsigned)i : (unsigned)(i - (d.nx - 2 * np)));
    const long nsz = (long)2 * np * d.pitch;
    const long rowz = (long)((j < np) ? j : j - (d.nz - 2 * np)) * d.pitch + i0;

    // levels s and s-1: f[s & 1] and f[(s & 1) ^ 1], selected from both sets
    // (already in SGPRs) rather than loaded by index
    const bool odd = s & 1;
    const long P = d.pitch;
    const long org = (long)(j + MNAD_ZO) * P + (i0 + MNAD_XO);   // storage of (j, i0)
    const gfloat* U = gp(odd ? d.f[1][0] : d.f[0][0]) + org;
    const gfloat* V = gp(odd ? d.f[1][1] : d.f[0][1]) + org;
    const gfloat* W = gp(odd ? d.f[1][2] : d.f[0][2]) + org;
    Nbh u, v, w;
    float vww, vee, wnn, wss;
    {
        const gfloat* a = (U - P) + lane;
        u.nw = a[-1]; u.n = a[0]; u.ne = a[1];
        a = U + lane;
        u.w = a[-1];  u.c = a[0]; u.e = a[1];
        a = (U + P) + lane;
        u.sw = a[-1]; u.s = a[0]; u.se = a[1];
        a = (V - P) + lane;
        v.nw = a[-1]; v.n = a[0]; v.ne = a[1];
        a = V + lane;
        vww = a[-2];  v.w = a[-1]; v.c = a[0]; v.e = a[1]; vee = a[2];
        a = (V + P) + lane;
        v.sw = a[-1]; v.s = a[0]; v.se = a[1];
        wnn = ((W - 2 * P) + lane)[0];
        a = (W - P) + lane;
        w.nw = a[-1]; w.n = a[0]; w.ne = a[1];
        a = W + lane;
        w.w = a[-1];  w.c = a[0]; w.e = a[1];
        a = (W + P) + lane;
        w.sw = a[-1]; w.s = a[0]; w.se = a[1];
        wss = ((W + 2 * P) + lane)[0];
    }
    const float rx = (gp(d.rx) + org)[lane];
    gfloat* PU = gp(odd ? d.f[0][0] : d.f[1][0]) + org;
    gfloat* PV = gp(odd ? d.f[0][1] : d.f[1][1]) + org;
    gfloat* PW = gp(odd ? d.f[0][2] : d.f[1][2]) + org;
    const float pu = PU[lane], pv = PV[lane], pw = PW[lane];

    int r = -1;
    if (d.nrec > 0) {
        const int off = __builtin_amdgcn_readfirstlane(d.rrow[j]);   // wave-uniform
        if (off >= 0) r = uni(gp(d.rcol) + off + i0)[lane];
    }

    const float rz = rx * d.kap;
    float du, dv, dw;
    mnad_incr(d.ic, rx, u, v, vww, vee, w, wnn, wss, du, dv, dw);

    if (inx || inz) {
        // The CPML memories see the pre-injection U (fwd_PDE_exact_MNAD.m:95-111).
        float uc = u.c, un = u.n, us = u.s, ue = u.e, uw = u.w;
        if (d.mode == MODE_FWD && d.src_j >= 0) {
            const float a0 = d.src_amp[s];
            if (j == d.src_j) {
                if (i == d.src_i) uc -= a0;
                if (i - 1 == d.src_i) uw -= a0;
                if (i + 1 == d.src_i) ue -= a0;
            }
            if (i == d.src_i) {
                if (j - 1 == d.src_j) un -= a0;
                if (j + 1 == d.src_j) us -= a0;
            }
        }
        if (inx) {
            PsiW mx;
            psi_load(mx, gp(d.psx) + rowx, gp(d.cfx) + rowx, nsx, bx);
            psi_step(mx, rx, uw, uc, ue, v.w, v.c, v.e, w.w, w.c, w.e, du, dv, dw);
            psi_store(mx, gp(d.psx) + rowx, nsx, bx);
        }
        if (inz) {
            PsiW mz;
            psi_load(mz, gp(d.psz) + rowz, gp(d.cfz) + rowz, nsz, 4u * lane);
            psi_step(mz, rz, un, uc, us, v.n, v.c, v.s, w.n, w.c, w.s, du, dv, dw);
            psi_store(mz, gp(d.psz) + rowz, nsz, 4u * lane);
        }
    }

    float nu = (2.0f * u.c + du) - pu;
    const float nv = (2.0f * v.c + dv) - pv;
    const float nw = (2.0f * w.c + dw) - pw;
    if (d.mode == MODE_FWD) {
        if (j == d.src_j && i == d.src_i) nu += d.src_amp[s + 1];
        if (r >= 0) (gp(d.gather) + (long)s * d.nrec)[(unsigned)r] = u.c;
    } else if (r >= 0) {
        nu -= (gp(d.radj) + (long)s * d.nrec)[(unsigned)r];
    }
    PU[lane] = nu;
    PV[lane] = nv;
    PW[lane] = nw;
    if (d.rord > 0 && (inx || inz)) {   // ring cells lie on the CPML strips
        const int ri = ring_index(j, i, d.nz, d.nx, np);
        if (ri >= 0) ring_put(d, s + 1, (unsigned)ri, nu, nv, nw);
    }

    if (d.T > 0 && !inx && !inz)
        snap_image_w(d, s, (long)(j - np) * (d.nx - 2 * np), (unsigned)(i - np), du, nu);
}

template <int BZR>
__global__ __launch_bounds__(BX * BZR) void k_stepw(const ShotDev* __restrict__ shots,
                                                    const TileRef* __restrict__ tiles,
                                                    const int* __restrict__ base, int rel)
{
    // One exit test after one round of descriptor loads (a padding entry reads
    // shot 0's descriptor and exits with the others): separate early exits
    // would make each descriptor field a dependent round of scalar loads.  The
    // test reads a field of each of the descriptor's first two cache lines.
    const TileRef tr = tiles[blockIdx.x];
    const ShotDev& d = shots[tr.shot < 0 ? 0 : tr.shot];
    const int s = *base + rel - d.t0;
    const int tj = tr.tile / d.ntx;
    const int ti = tr.tile - tj * d.ntx;
    const int j = tj * BZR + __builtin_amdgcn_readfirstlane((int)threadIdx.y);   // wave row
    if ((tr.shot < 0) | (s < 0) | (s >= d.nt) | (j >= d.nz) | (d.rx == nullptr)) return;
    if (d.mode == MODE_REC) rec_one<false>(d, s, j, ti * BX, threadIdx.x);
    else step_one<false>(d, s, j, ti * BX, threadIdx.x);
}

// ------------------------------------------------ two rows per wavefront (kernel 11)
// A wavefront owns 64 columns of rows j and j+1.  The rows the two updates
// share are loaded once: U and Vs rows j-1 .. j+2, Ws rows j-2 .. j+3 (the
// outer two at the centre column only), 14 row loads for two points against
// 22 for two one-row waves, and the two increments' arithmetic interleaves.
// The tail of each point (CPML memories, injections, receivers, stores,
// snapshots / imaging) is step_one's.

__device__ __forceinline__ void pair_tail(const ShotDev& d, int s, int j, int i0, unsigned lane,
                                          const Nbh& u, const Nbh& v, const Nbh& w, float rx,
                                          float pu, float pv, float pw, float du, float dv,
                                          float dw, gfloat* PU, gfloat* PV, gfloat* PW)
{
    const int i = i0 + (int)lane;
    const int np = d.npml;
    const bool inx = (i < np) || (i >= d.nx - np);
    const bool inz = (j < np) || (j >= d.nz - np);   // wave-uniform
    int r = -1;
    if (d.nrec > 0) {
        const int off = __builtin_amdgcn_readfirstlane(d.rrow[j]);   // wave-uniform
        if (off >= 0) r = uni(gp(d.rcol) + off + i0)[lane];
    }
    if (inx || inz) {
        // The CPML memories see the pre-injection U (fwd_PDE_exact_MNAD.m:95-111).
        float uc = u.c, un = u.n, us = u.s, ue = u.e, uw = u.w;
        if (d.mode == MODE_FWD && d.src_j >= 0) {
            const float a0 = d.src_amp[s];
            if (j == d.src_j) {
                if (i == d.src_i) uc -= a0;
                if (i - 1 == d.src_i) uw -= a0;
                if (i + 1 == d.src_i) ue -= a0;
            }
            if (i == d.src_i) {
                if (j - 1 == d.src_j) un -= a0;
                if (j + 1 == d.src_j) us -= a0;
            }
        }
        if (inx) {
            const long nsx = (long)d.nz * 2 * np, rowx = (long)j * 2 * np;
            const unsigned bx = 4u * ((i < np) ? (unsigned)i : (unsigned)(i - (d.nx - 2 * np)));
            PsiW mx;
            psi_load(mx, gp(d.psx) + rowx, gp(d.cfx) + rowx, nsx, bx);
            psi_step(mx, rx, uw, uc, ue, v.w, v.c, v.e, w.w, w.c, w.e, du, dv, dw);
            psi_store(mx, gp(d.psx) + rowx, nsx, bx);
        }
        if (inz) {
            const long nsz = (long)2 * np * d.pitch;
            const long rowz = (long)((j < np) ? j : j - (d.nz - 2 * np)) * d.pitch + i0;
            PsiW mz;
            psi_load(mz, gp(d.psz) + rowz, gp(d.cfz) + rowz, nsz, 4u * lane);
            psi_step(mz, rx * d.kap, un, uc, us, v.n, v.c, v.s, w.n, w.c, w.s, du, dv, dw);
            psi_store(mz, gp(d.psz) + rowz, nsz, 4u * lane);
        }
    }
    float nu = (2.0f * u.c + du) - pu;
    const float nv = (2.0f * v.c + dv) - pv;
    const float nw = (2.0f * w.c + dw) - pw;
    if (d.mode == MODE_FWD) {
        if (j == d.src_j && i == d.src_i) nu += d.src_amp[s + 1];
        if (r >= 0) (gp(d.gather) + (long)s * d.nrec)[(unsigned)r] = u.c;
    } else if (r >= 0) {
        nu -= (gp(d.radj) + (long)s * d.nrec)[(unsigned)r];
    }
    PU[lane] = nu;
    PV[lane] = nv;
    PW[lane] = nw;
    if (d.rord > 0 && (inx || inz)) {
        const int ri = ring_index(j, i, d.nz, d.nx, np);
        if (ri >= 0) ring_put(d, s + 1, (unsigned)ri, nu, nv, nw);
    }
    if (d.T > 0 && !inx && !inz)
        snap_image_w(d, s, (long)(j - np) * (d.nx - 2 * np), (unsigned)(i - np), du, nu);
}

__device__ __forceinline__ void step_pair(const ShotDev& d, int s, int j, int i0, unsigned lane)
{
    const int i = i0 + (int)lane;
    if (i >= d.nx) return;
    const bool two = j + 1 < d.nz;   // wave-uniform; rows past the grid read zero ghost rows
    const bool odd = s & 1;
    const long P = d.pitch;
    const long org = (long)(j + MNAD_ZO) * P + (i0 + MNAD_XO);   // storage of (j, i0)
    const gfloat* U = gp(odd ? d.f[1][0] : d.f[0][0]) + org + lane;
    const gfloat* V = gp(odd ? d.f[1][1] : d.f[0][1]) + org + lane;
    const gfloat* W = gp(odd ? d.f[1][2] : d.f[0][2]) + org + lane;
    // U rows j-1 .. j+2 and Vs rows j-1 .. j+2 at columns -1 .. +1 (Vs rows j, j+1
    // at -2 .. +2), Ws rows j-1 .. j+2 at -1 .. +1 and rows j-2, j+3 at 0
    const float u0w = U[-P - 1], u0c = U[-P], u0e = U[-P + 1];
    const float u1w = U[-1], u1c = U[0], u1e = U[1];
    const float u2w = U[P - 1], u2c = U[P], u2e = U[P + 1];
    const float u3w = U[2 * P - 1], u3c = U[2 * P], u3e = U[2 * P + 1];
    const float v0w = V[-P - 1], v0c = V[-P], v0e = V[-P + 1];
    const float v1ww = V[-2], v1w = V[-1], v1c = V[0], v1e = V[1], v1ee = V[2];
    const float v2ww = V[P - 2], v2w = V[P - 1], v2c = V[P], v2e = V[P + 1], v2ee = V[P + 2];
    const float v3w = V[2 * P - 1], v3c = V[2 * P], v3e = V[2 * P + 1];
    const float wm2 = W[-2 * P];
    const float w0w = W[-P - 1], w0c = W[-P], w0e = W[-P + 1];
    const float w1w = W[-1], w1c = W[0], w1e = W[1];
    const float w2w = W[P - 1], w2c = W[P], w2e = W[P + 1];
    const float w3w = W[2 * P - 1], w3c = W[2 * P], w3e = W[2 * P + 1];
    const float wp3 = W[3 * P];
    const gfloat* RX = gp(d.rx) + org + lane;
    const float rxa = RX[0], rxb = RX[P];
    gfloat* PU = gp(odd ? d.f[0][0] : d.f[1][0]) + org;
    gfloat* PV = gp(odd ? d.f[0][1] : d.f[1][1]) + org;
    gfloat* PW = gp(odd ? d.f[0][2] : d.f[1][2]) + org;
    const float pua = PU[lane], pva = PV[lane], pwa = PW[lane];
    const float pub = (PU + P)[lane], pvb = (PV + P)[lane], pwb = (PW + P)[lane];

    Nbh ua, va, wa, ub, vb, wb;
    ua.nw = u0w; ua.n = u0c; ua.ne = u0e; ua.w = u1w; ua.c = u1c; ua.e = u1e;
    ua.sw = u2w; ua.s = u2c; ua.se = u2e;
    ub.nw = u1w; ub.n = u1c; ub.ne = u1e; ub.w = u2w; ub.c = u2c; ub.e = u2e;
    ub.sw = u3w; ub.s = u3c; ub.se = u3e;
    va.nw = v0w; va.n = v0c; va.ne = v0e; va.w = v1w; va.c = v1c; va.e = v1e;
    va.sw = v2w; va.s = v2c; va.se = v2e;
    vb.nw = v1w; vb.n = v1c; vb.ne = v1e; vb.w = v2w; vb.c = v2c; vb.e = v2e;
    vb.sw = v3w; vb.s = v3c; vb.se = v3e;
    wa.nw = w0w; wa.n = w0c; wa.ne = w0e; wa.w = w1w; wa.c = w1c; wa.e = w1e;
    wa.sw = w2w; wa.s = w2c; wa.se = w2e;
    wb.nw = w1w; wb.n = w1c; wb.ne = w1e; wb.w = w2w; wb.c = w2c; wb.e = w2e;
    wb.sw = w3w; wb.s = w3c; wb.se = w3e;
    float dua, dva, dwa, dub, dvb, dwb;
    mnad_incr(d.ic, rxa, ua, va, v1ww, v1ee, wa, wm2, w3c, dua, dva, dwa);
    mnad_incr(d.ic, rxb, ub, vb, v2ww, v2ee, wb, w0c, wp3, dub, dvb, dwb);

    pair_tail(d, s, j, i0, lane, ua, va, wa, rxa, pua, pva, pwa, dua, dva, dwa, PU, PV, PW);
    if (two)
        pair_tail(d, s, j + 1, i0, lane, ub, vb, wb, rxb, pub, pvb, pwb, dub, dvb, dwb, PU + P,
                  PV + P, PW + P);
}

__global__ __launch_bounds__(BX) __attribute__((amdgpu_waves_per_eu(6))) void k_stepw2(const ShotDev* __restrict__ shots,
                                               const TileRef* __restrict__ tiles,
                                               const int* __restrict__ base, int rel)
{
    // one exit test after one round of descriptor loads, as in k_stepw
    const TileRef tr = tiles[blockIdx.x];
    const ShotDev& d = shots[tr.shot < 0 ? 0 : tr.shot];
    const int s = *base + rel - d.t0;
    const int tj = tr.tile / d.ntx;
    const int ti = tr.tile - tj * d.ntx;
    const int j = 2 * tj;
    if ((tr.shot < 0) | (s < 0) | (s >= d.nt) | (j >= d.nz) | (d.rx == nullptr)) return;
    if (d.mode == MODE_REC) {   // reverse reconstruction: one row at a time
        rec_one<false>(d, s, j, ti * BX, threadIdx.x);
        if (j + 1 < d.nz) rec_one<false>(d, s, j + 1, ti * BX, threadIdx.x);
        return;
    }
    step_pair(d, s, j, ti * BX, threadIdx.x);
}

// ------------------------------------------------ resident blocks (kernel 9)
// One cooperative launch runs every step of a batch.  Workgroup b owns
// blocks[b] for the whole launch: columns c0 .. c0+63 of one shot and rows
// r0 .. r0+bh-1.
//  - Level s of U, Vs, Ws sits in LDS with a ring of the two cells around the
//    block, in two buffers: step s reads one and writes level s+1 into the
//    other, over level s-1, which is the update's previous level at the same
//    point.  rx sits in a third LDS plane.  The per-point update is
//    step_one's, read from LDS.
//  - Each step computes the edge cells first (rows 0, 1, bh-2, bh-1 and
//    columns 0, 1, 62, 63: waves 0-3 take the four edge rows, waves 4-7 one
//    edge column each, a lane per row) and sends them out at once, then the
//    interior (rows 2 .. bh-3, columns 2 .. 61, a wave per row), behind which
//    the neighbours' edges arrive.
//  - Edge cells go out as 64-bit records {value, level} in the block's edge
//    area of `ebuf`, one agent-scope atomic store each; level L uses slot
//    L & 1.  The ring of level s comes back as agent-scope atomic loads of the
//    neighbours' records, each retried until it carries level s.  Both are
//    coherent per access across the XCDs: no flags, fences or cache-wide
//    write-back or invalidation.
//  - Two slots suffice: a block writes level L+1 (over level L-1) only after
//    it has read its neighbours' level L, which they publish after they have
//    read its level L-1.
//  - Every wait is bounded (`spin` polls of a record): a block that gives up
//    sets *err and leaves, and blocks that see *err leave too, so the launch
//    always drains.  The launch is cooperative, so all blocks are resident at
//    once.
//  - After its last step a block writes its last two levels in full to the
//    level buffers in global memory.
// CPML memories, gathers, snapshots and the image stay in global memory; each
// point's are touched only by the block that owns it.
constexpr int RES_NW = 8;            // wavefronts per block
constexpr int RES_LW = BX + 4;       // LDS row: the block's 64 columns and 2 each side

// Edge area of one block: [slot][field][ne] records, ne = 256 + 4 bh:
//   top rows 0, 1 (r * 64 + c), bottom rows bh-2, bh-1 (128 + ...),
//   left columns 0, 1 (256 + 2 r + c), right columns 62, 63 (256 + 2 bh + ...).
__host__ __device__ __forceinline__ int res_edge_cells(int bh) { return 256 + 4 * bh; }

typedef unsigned long long u64;

__device__ __forceinline__ u64 edge_rec(float x, int level)
{
    return (u64)__float_as_uint(x) | ((u64)(unsigned)level << 32);
}

// What every point of a block's step shares.
struct ResStep {
    const float* LU;   // level s, buffer with ring
    float* NU;         // level s+1 written over level s-1
    const float* LRX;  // rx [bh x 64]
    u64* out;          // this block's records of level s+1
    int plane, ne, bh, r0, c0, s;
    float a0, a1;      // source samples s and s+1
    bool src;
};

// One point (block row r, column c; grid row j, column i) of step s.  ROWW:
// the wave holds one block row (j wave-uniform, row origins in SGPRs);
// otherwise a lane per row (the edge-column waves).  rr: receiver slot or -1.
template <bool ROWW>
__device__ __forceinline__ void res_point(const ShotDev& d, const ResStep& t, int r, int c,
                                          int rr)
{
    const int j = t.r0 + r, i = t.c0 + c;
    const int np = d.npml, nz = d.nz, nx = d.nx;
    const long P = d.pitch;
    const int o = (r + 2) * RES_LW + (c + 2);
    const float* LU = t.LU;
    const float* LV = LU + t.plane;
    const float* LW = LU + 2 * t.plane;
    Nbh u, v, w;
    u.nw = LU[o - RES_LW - 1]; u.n = LU[o - RES_LW]; u.ne = LU[o - RES_LW + 1];
    u.w = LU[o - 1];           u.c = LU[o];          u.e = LU[o + 1];
    u.sw = LU[o + RES_LW - 1]; u.s = LU[o + RES_LW]; u.se = LU[o + RES_LW + 1];
    v.nw = LV[o - RES_LW - 1]; v.n = LV[o - RES_LW]; v.ne = LV[o - RES_LW + 1];
    v.w = LV[o - 1];           v.c = LV[o];          v.e = LV[o + 1];
    v.sw = LV[o + RES_LW - 1]; v.s = LV[o + RES_LW]; v.se = LV[o + RES_LW + 1];
    w.nw = LW[o - RES_LW - 1]; w.n = LW[o - RES_LW]; w.ne = LW[o - RES_LW + 1];
    w.w = LW[o - 1];           w.c = LW[o];          w.e = LW[o + 1];
    w.sw = LW[o + RES_LW - 1]; w.s = LW[o + RES_LW]; w.se = LW[o + RES_LW + 1];
    const float vww = LV[o - 2], vee = LV[o + 2];
    const float wnn = LW[o - 2 * RES_LW], wss = LW[o + 2 * RES_LW];
    const float rx = t.LRX[r * BX + c], rz = rx * d.kap;
    float du, dv, dw;
    mnad_incr(d.ic, rx, u, v, vww, vee, w, wnn, wss, du, dv, dw);

    const bool inx = (i < np) || (i >= nx - np);
    const bool inz = (j < np) || (j >= nz - np);
    if (inx || inz) {
        // The CPML memories see the pre-injection U (fwd_PDE_exact_MNAD.m:95-111).
        float uc = u.c, un = u.n, us = u.s, ue = u.e, uw = u.w;
        if (t.src) {
            const float a0 = t.a0;
            if (j == d.src_j) {
                if (i == d.src_i) uc -= a0;
                if (i - 1 == d.src_i) uw -= a0;
                if (i + 1 == d.src_i) ue -= a0;
            }
            if (i == d.src_i) {
                if (j - 1 == d.src_j) un -= a0;
                if (j + 1 == d.src_j) us -= a0;
            }
        }
        if (inx) {
            const long nsx = (long)nz * 2 * np;
            const unsigned xo = (i < np) ? (unsigned)i : (unsigned)(i - (nx - 2 * np));
            PsiW mx;
            if (ROWW) {
                const long rowx = (long)j * 2 * np;
                psi_load(mx, gp(d.psx) + rowx, gp(d.cfx) + rowx, nsx, 4u * xo);
                psi_step(mx, rx, uw, uc, ue, v.w, v.c, v.e, w.w, w.c, w.e, du, dv, dw);
                psi_store(mx, gp(d.psx) + rowx, nsx, 4u * xo);
            } else {
                const unsigned b = 4u * ((unsigned)j * 2u * (unsigned)np + xo);
                psi_load(mx, gp(d.psx), gp(d.cfx), nsx, b);
                psi_step(mx, rx, uw, uc, ue, v.w, v.c, v.e, w.w, w.c, w.e, du, dv, dw);
                psi_store(mx, gp(d.psx), nsx, b);
            }
        }
        if (inz) {
            const long nsz = (long)2 * np * P;
            const int zr = (j < np) ? j : j - (nz - 2 * np);
            PsiW mz;
            if (ROWW) {
                const long rowz = (long)zr * P + t.c0;
                psi_load(mz, gp(d.psz) + rowz, gp(d.cfz) + rowz, nsz, 4u * (unsigned)c);
                psi_step(mz, rz, un, uc, us, v.n, v.c, v.s, w.n, w.c, w.s, du, dv, dw);
                psi_store(mz, gp(d.psz) + rowz, nsz, 4u * (unsigned)c);
            } else {
                const unsigned b = 4u * ((unsigned)zr * (unsigned)P + (unsigned)c);
                psi_load(mz, gp(d.psz) + t.c0, gp(d.cfz) + t.c0, nsz, b);
                psi_step(mz, rz, un, uc, us, v.n, v.c, v.s, w.n, w.c, w.s, du, dv, dw);
                psi_store(mz, gp(d.psz) + t.c0, nsz, b);
            }
        }
    }

    float* NU = t.NU;
    // level s-1 at this point: the cell level s+1 replaces
    float nu = (2.0f * u.c + du) - NU[o];
    const float nv = (2.0f * v.c + dv) - NU[t.plane + o];
    const float nw = (2.0f * w.c + dw) - NU[2 * t.plane + o];
    const int s = t.s;
    if (d.mode == MODE_FWD) {
        if (j == d.src_j && i == d.src_i) nu += t.a1;
        if (rr >= 0) (gp(d.gather) + (long)s * d.nrec)[(unsigned)rr] = u.c;
    } else if (rr >= 0) {
        nu -= (gp(d.radj) + (long)s * d.nrec)[(unsigned)rr];
    }
    NU[o] = nu;
    NU[t.plane + o] = nv;
    NU[2 * t.plane + o] = nw;
    // edge records of level s+1 (a corner cell goes to two of them)
    const int bh = t.bh;
    int e0 = -1, e1 = -1;
    if (r < 2) e0 = r * BX + c;
    else if (r >= bh - 2) e0 = 128 + (r - (bh - 2)) * BX + c;
    if (c < 2) e1 = 256 + 2 * r + c;
    else if (c >= BX - 2) e1 = 256 + 2 * bh + 2 * r + (c - (BX - 2));
    if (e0 >= 0 || e1 >= 0) {
        const float nf[3] = {nu, nv, nw};
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            const u64 x = edge_rec(nf[f], s + 1);
            if (e0 >= 0)
                __hip_atomic_store(t.out + f * t.ne + e0, x, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            if (e1 >= 0)
                __hip_atomic_store(t.out + f * t.ne + e1, x, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    if (d.T > 0 && !inx && !inz)
        snap_image_w(d, s, (long)(j - np) * (nx - 2 * np), (unsigned)(i - np), du, nu);
}

// Receiver slot of grid point (j, i), -1 if none.  The load is consumed
// before returning, so no load stays in flight past the call.
__device__ __forceinline__ int res_rec(const ShotDev& d, int j, int i)
{
    if (d.nrec <= 0) return -1;
    const int off = d.rrow[j];
    int x = off >= 0 ? d.rcol[off + i] : -1;
    asm volatile("" : "+v"(x));
    return x;
}

__global__ __launch_bounds__(BX * RES_NW) void k_res(const ShotDev* __restrict__ shots,
                                                     const ResBlock* __restrict__ blocks,
                                                     u64* ebuf, int* err, int nsteps, int bh,
                                                     int spin)
{
    // spin: polls of one edge record before the wait gives up
    // [2 buffers][U, Vs, Ws][(bh + 4) x RES_LW], then rx [bh x BX]
    extern __shared__ float lds[];
    __shared__ int give_up;
    const ResBlock& rb = blocks[blockIdx.x];
    const ShotDev& d = shots[rb.shot];
    const int lane = (int)threadIdx.x;
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int tid = wv * BX + lane;
    const int nst = min(nsteps, d.nt);
    if (nst <= 0) return;
    const int plane = (bh + 4) * RES_LW;
    float* LRX = lds + 6 * plane;
    const int ne = res_edge_cells(bh);
    const long area = 2L * 3 * ne;   // records per block
    const int nz = d.nz, nx = d.nx;
    const long P = d.pitch;
    const int r0 = rb.r0, c0 = rb.c0;
    const int ring = 4 * RES_LW + 4 * bh;   // cells of the ring
    u64* mine = ebuf + (long)blockIdx.x * area;
    if (tid == 0) give_up = 0;

    // Level 0 into buffer 0, level -1 into buffer 1, rx.  Cells past the grid
    // (a partial block's last rows or columns) stand for the zero ghost cells
    // and are never written again, so they start cleared in both buffers.
    for (int t = tid; t < (bh + 4) * RES_LW; t += BX * RES_NW) {
        const int rr = t / RES_LW - 2, cc = t % RES_LW - 2;
        const int j = r0 + rr, ii = c0 + cc;
        const bool in = j >= 0 && j < nz && ii >= 0 && ii < nx;
        const long o = (long)(j + MNAD_ZO) * P + (ii + MNAD_XO);
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            lds[f * plane + t] = in ? d.f[0][f][o] : 0.0f;
            lds[(3 + f) * plane + t] = in ? d.f[1][f][o] : 0.0f;
        }
        if (rr >= 0 && rr < bh && cc >= 0 && cc < BX) LRX[rr * BX + cc] = in ? d.rx[o] : 0.0f;
    }

    // This wave's points.  Edge pass: waves 0-3 block row erow (a lane per
    // column), waves 4-7 block column ecol of rows 2 .. bh-3 (a lane per row).
    // Interior pass: rows 2 + wv, 2 + wv + RES_NW, ... (lanes 2 .. 61).
    const int erow = wv == 0 ? 0 : (wv == 1 ? 1 : (wv == 2 ? bh - 2 : bh - 1));
    const int ecol = wv == 4 ? 0 : (wv == 5 ? 1 : (wv == 6 ? BX - 2 : BX - 1));
    const bool ecol_on = wv >= 4 && lane < bh - 4 && r0 + 2 + lane < nz && c0 + ecol < nx;
    const bool erow_on = wv < 4 && r0 + erow < nz && c0 + lane < nx;
    // receiver slots: the edge point's (looked up once), and the first two
    // interior rows with receivers (in registers; a global load inside the
    // row loop would make every row wait for the stores before it); further
    // receiver rows look theirs up each step
    int rec_e = -1;
    if (ecol_on) rec_e = res_rec(d, r0 + 2 + lane, c0 + ecol);
    if (erow_on) rec_e = res_rec(d, r0 + erow, c0 + lane);
    int ra = -1, rb2 = -1, reca = -1, recb = -1;
    bool more = false;
    if (d.nrec > 0) {
        for (int r = 2 + wv; r < bh - 2 && r0 + r < nz; r += RES_NW) {
            const int off = __builtin_amdgcn_readfirstlane(d.rrow[r0 + r]);
            if (off < 0) continue;
            const int x = c0 + lane < nx ? d.rcol[off + c0 + lane] : -1;
            if (ra < 0) {
                ra = r;
                reca = x;
            } else if (rb2 < 0) {
                rb2 = r;
                recb = x;
            } else {
                more = true;
            }
        }
    }

    // The ring cells this thread reloads every step (at most two: the ring
    // has at most 4 * 68 + 4 * 64 cells for 512 threads), as LDS offsets and
    // the neighbour records that hold them (-1: past the grid, stays zero).
    int rl[2];
    long rsrc[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int t = tid + h * BX * RES_NW;
        rl[h] = -1;
        rsrc[h] = -1;
        if (t >= ring) continue;
        int rr, cc;
        if (t < 4 * RES_LW) {
            const int k = t / RES_LW;
            rr = k < 2 ? k - 2 : bh + k - 2;
            cc = t % RES_LW - 2;
        } else {
            const int u = t - 4 * RES_LW, k = u & 3;
            rr = u >> 2;
            cc = k < 2 ? k - 2 : BX + k - 2;
        }
        rl[h] = (rr + 2) * RES_LW + (cc + 2);
        const int j = r0 + rr, ii = c0 + cc;
        if (j < 0 || j >= nz || ii < 0 || ii >= nx) continue;
        // owner block (oz, ox) relative to this one, and the cell's record there
        const int oz = rr < 0 ? -1 : (rr >= bh ? 1 : 0);
        const int ox = cc < 0 ? -1 : (cc >= BX ? 1 : 0);
        const int lr = rr - oz * bh, lc = cc - ox * BX;   // row / column in the owner
        int idx;
        if (oz < 0) idx = 128 + (lr - (bh - 2)) * BX + lc;   // owner's bottom rows
        else if (oz > 0) idx = lr * BX + lc;                 // owner's top rows
        else if (ox < 0) idx = 256 + 2 * bh + 2 * lr + (lc - (BX - 2));   // owner's right columns
        else idx = 256 + 2 * lr + lc;                        // owner's left columns
        const int k = (oz + 1) * 3 + (ox + 1);
        const int nb = rb.nbr[k < 4 ? k : k - 1];
        if (nb >= 0) rsrc[h] = (long)nb * area + idx;
    }
    __syncthreads();

    ResStep st;
    st.plane = plane;
    st.ne = ne;
    st.bh = bh;
    st.r0 = r0;
    st.c0 = c0;
    st.LRX = LRX;
    st.src = d.mode == MODE_FWD && d.src_j >= 0;
    int cur = 0;   // buffer holding level s
    for (int s = 0; s < nst; ++s) {
        // this step's source samples, loaded before the barrier so that no
        // load is left in flight inside the row loops
        st.a0 = st.a1 = 0.0f;
        if (st.src) {
            st.a0 = d.src_amp[s];
            st.a1 = d.src_amp[s + 1];
        }
        __syncthreads();   // level s and its ring are in buffer cur
        if (give_up) return;

        st.s = s;
        st.LU = lds + cur * 3 * plane;
        st.NU = lds + (cur ^ 1) * 3 * plane;
        st.out = mine + (long)((s + 1) & 1) * 3 * ne;
        // edge pass: level s+1 of the edge cells goes out first
        if (erow_on) res_point<true>(d, st, erow, lane, rec_e);
        if (ecol_on) res_point<false>(d, st, 2 + lane, ecol, rec_e);
        // the ring of level s+1 (the neighbours' slot (s+1) & 1) is requested
        // now and arrives behind the interior pass
        const bool next = s + 1 < nst;
        const long so = (long)((s + 1) & 1) * 3 * ne;
        u64 x[2][3];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int f = 0; f < 3; ++f)
                x[h][f] = (next && rsrc[h] >= 0)
                              ? __hip_atomic_load(ebuf + rsrc[h] + so + f * ne, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT)
                              : 0ull;
        // interior pass
        const bool lin = lane >= 2 && lane < BX - 2 && c0 + lane < nx;
        for (int r = 2 + wv; r < bh - 2; r += RES_NW) {
            if (r0 + r >= nz) break;
            int rr = r == ra ? reca : (r == rb2 ? recb : -1);
            if (more && r != ra && r != rb2 && lin) rr = res_rec(d, r0 + r, c0 + lane);
            if (lin) res_point<true>(d, st, r, lane, rr);
        }
        // ring of level s+1 into the other buffer (nobody reads its ring
        // during step s); records not yet at level s+1 are polled
        if (next) {
            float* L = lds + (cur ^ 1) * 3 * plane;
            bool ok = true;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (rl[h] < 0) continue;
                if (rsrc[h] < 0) {
#pragma unroll
                    for (int f = 0; f < 3; ++f) L[f * plane + rl[h]] = 0.0f;
                    continue;
                }
                const u64* srcp = ebuf + rsrc[h] + so;
#pragma unroll
                for (int f = 0; f < 3; ++f) {
                    u64 y = x[h][f];
                    for (int n = 0; ok && (int)(y >> 32) != s + 1; ++n) {
                        if (n >= spin || __hip_atomic_load(err, __ATOMIC_RELAXED,
                                                               __HIP_MEMORY_SCOPE_AGENT) != 0) {
                            ok = false;
                            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                        y = __hip_atomic_load(srcp + f * ne, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
                    }
                    L[f * plane + rl[h]] = __uint_as_float((unsigned)y);
                }
            }
            if (!ok) give_up = 1;
        }
        cur ^= 1;
    }

    // last two levels in full: level nst in buffer cur, level nst-1 in the
    // other (other waves computed some of this wave's rows in the last step)
    __syncthreads();
    const float* LA = lds + cur * 3 * plane;
    const float* LB = lds + (cur ^ 1) * 3 * plane;
    const int i = c0 + lane;
    for (int r = wv; r < bh; r += RES_NW) {
        const int j = r0 + r;
        if (j >= nz || i >= nx) continue;
        const int o = (r + 2) * RES_LW + (lane + 2);
        const long g = (long)(j + MNAD_ZO) * P + (i + MNAD_XO);
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            d.f[nst & 1][f][g] = LA[f * plane + o];
            d.f[(nst + 1) & 1][f][g] = LB[f * plane + o];
        }
    }
}

// ------------------------------------------------ four (or two) columns per lane
// The interior of a shot (rows npml .. nz-npml-1, columns npml .. nx-npml-1:
// no CPML there) in waves of 64 lanes x NC columns (kernel 6: NC = 4,
// kernel 10: NC = 2).  Lane l holds columns i0 + NC l .. i0 + NC l + NC-1 of
// row j, one 4NC-byte load per field row, and takes the columns just left /
// right of its own from lanes l-1 / l+1 with DPP shifts across the wavefront
// (wave_shr:1 / wave_shl:1).  Lanes 0 and 63 only supply those halo columns,
// so a wave updates 62 NC columns with 15 loads and 3 stores per lane,
// against 64 columns with 15 loads for one point per lane.  All 64 lanes stay
// active through the loads and the shifts.  Lane 0's first column (i0) is a
// multiple of NC, so every row load is aligned to its size.
constexpr int X4W = 4 * (BX - 2);   // columns updated per wave, four columns per lane

// bound_ctrl set: a lane whose source lane is outside the wave reads 0, so
// the shift needs no preloaded "old" value (one instruction, and the DPP can
// fold into the instruction that uses it)
__device__ __forceinline__ float lane_prev(float x)   // lane l-1's x (0 in lane 0)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x),
                                                                 0x138, 0xf, 0xf, true));
}

__device__ __forceinline__ float lane_next(float x)   // lane l+1's x (0 in lane 63)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x),
                                                                 0x130, 0xf, 0xf, true));
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) f32x4 gf32x4;
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) f32x2 gf32x2;

template <int NC> struct VecN;
template <> struct VecN<4> { typedef f32x4 T; typedef gf32x4 G; };
template <> struct VecN<2> { typedef f32x2 T; typedef gf32x2 G; };

// Columns -1 .. NC of a lane's NC.
template <int NC>
struct RowN {
    float e[NC], l, r;
    __device__ __forceinline__ float at(int c) const { return c < 0 ? l : (c > NC - 1 ? r : e[c]); }
};

template <int NC>
__device__ __forceinline__ RowN<NC> rown(const gfloat* base, unsigned lane)
{
    const typename VecN<NC>::T x = ((const typename VecN<NC>::G*)base)[lane];
    RowN<NC> o;
#pragma unroll
    for (int e = 0; e < NC; ++e) o.e[e] = x[e];
    o.l = lane_prev(x[NC - 1]);
    o.r = lane_next(x[0]);
    return o;
}

template <int NC>
__device__ __forceinline__ void step_xn(const ShotDev& d, int s, int j, int i0, int lo, int hi,
                                        unsigned lane)
{
    typedef typename VecN<NC>::T VT;
    typedef typename VecN<NC>::G GVT;
    // current level: s, or ntf-1-s in a reverse reconstruction (rec_one)
    const bool rec = d.mode == MODE_REC;
    const int lev = rec ? d.ntf - 1 - s : s;
    const int cb = lev & 1;
    const long P = d.pitch;
    const long org = (long)(j + MNAD_ZO) * P + (i0 + MNAD_XO);   // storage of (j, i0)
    const gfloat* U = gp(d.f[cb][0]) + org;
    const gfloat* V = gp(d.f[cb][1]) + org;
    const gfloat* W = gp(d.f[cb][2]) + org;
    const RowN<NC> un = rown<NC>(U - P, lane), uc = rown<NC>(U, lane), us = rown<NC>(U + P, lane);
    const RowN<NC> vn = rown<NC>(V - P, lane), vc = rown<NC>(V, lane), vs = rown<NC>(V + P, lane);
    const RowN<NC> wn = rown<NC>(W - P, lane), wc = rown<NC>(W, lane), ws = rown<NC>(W + P, lane);
    const VT wnn = ((const GVT*)(W - 2 * P))[lane];
    const VT wss = ((const GVT*)(W + 2 * P))[lane];
    // columns -2 and NC+1 (the Vs row's second neighbours beyond the lane)
    const float vl2 = lane_prev(vc.e[NC - 2]), vr2 = lane_next(vc.e[1]);
    const VT rx4 = ((const GVT*)(gp(d.rx) + org))[lane];
    float nu[NC], nv[NC], nw[NC], dus[NC];
#pragma unroll
    for (int e = 0; e < NC; ++e) {
        Nbh u, v, w;
        u.nw = un.at(e - 1); u.n = un.e[e]; u.ne = un.at(e + 1);
        u.w = uc.at(e - 1);  u.c = uc.e[e]; u.e = uc.at(e + 1);
        u.sw = us.at(e - 1); u.s = us.e[e]; u.se = us.at(e + 1);
        v.nw = vn.at(e - 1); v.n = vn.e[e]; v.ne = vn.at(e + 1);
        v.w = vc.at(e - 1);  v.c = vc.e[e]; v.e = vc.at(e + 1);
        v.sw = vs.at(e - 1); v.s = vs.e[e]; v.se = vs.at(e + 1);
        w.nw = wn.at(e - 1); w.n = wn.e[e]; w.ne = wn.at(e + 1);
        w.w = wc.at(e - 1);  w.c = wc.e[e]; w.e = wc.at(e + 1);
        w.sw = ws.at(e - 1); w.s = ws.e[e]; w.se = ws.at(e + 1);
        const float vww = e >= 2 ? vc.e[e - 2] : (e == 1 ? vc.l : vl2);
        const float vee = e <= NC - 3 ? vc.e[e + 2] : (e == NC - 2 ? vc.r : vr2);
        const float rx = rx4[e];
        float du, dv, dw;
        mnad_incr(d.ic, rx, u, v, vww, vee, w, wnn[e], wss[e], du, dv, dw);
        nu[e] = 2.0f * u.c + du;
        nv[e] = 2.0f * v.c + dv;
        nw[e] = 2.0f * w.c + dw;
        dus[e] = du;
        // one column at a time: interleaving them raises the register count
        if (NC == 4) __builtin_amdgcn_sched_barrier(0);
    }
    // the previous level, loaded after the arithmetic (fewer live registers)
    GVT* PU = (GVT*)(gp(d.f[cb ^ 1][0]) + org);
    GVT* PV = (GVT*)(gp(d.f[cb ^ 1][1]) + org);
    GVT* PW = (GVT*)(gp(d.f[cb ^ 1][2]) + org);
    {
        const VT pu = PU[lane], pv = PV[lane], pw = PW[lane];
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            nu[e] -= pu[e];
            nv[e] -= pv[e];
            nw[e] -= pw[e];
        }
    }

    // owned columns: [lo, hi) of this tile, lanes 1 .. 62
    const int c0 = i0 + NC * (int)lane;
    if (lane < 1 || lane > BX - 2 || c0 + NC - 1 < lo || c0 >= hi) return;
    if ((d.mode == MODE_FWD || rec) && j == d.src_j) {
#pragma unroll
        for (int e = 0; e < NC; ++e)
            if (c0 + e == d.src_i) nu[e] += d.src_amp[lev + 1];
    }
    if (d.nrec > 0) {
        const int off = __builtin_amdgcn_readfirstlane(d.rrow[j]);   // wave-uniform
        if (off >= 0) {
            const gint* rc = uni(gp(d.rcol) + off);
#pragma unroll
            for (int e = 0; e < NC; ++e) {
                const int c = c0 + e;
                if (c < lo || c >= hi) continue;
                const int r = rc[(unsigned)c];
                if (r < 0) continue;
                if (d.mode == MODE_FWD)
                    (gp(d.gather) + (long)s * d.nrec)[(unsigned)r] = uc.e[e];
                else
                    nu[e] -= (gp(d.radj) + (long)s * d.nrec)[(unsigned)r];
            }
        }
    }
    if (c0 >= lo && c0 + NC - 1 < hi) {
        VT a, b2, c2;
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            a[e] = nu[e];
            b2[e] = nv[e];
            c2[e] = nw[e];
        }
        PU[lane] = a;
        PV[lane] = b2;
        PW[lane] = c2;
    } else {
        gfloat* qu = (gfloat*)(PU + lane);
        gfloat* qv = (gfloat*)(PV + lane);
        gfloat* qw = (gfloat*)(PW + lane);
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            if (c0 + e < lo || c0 + e >= hi) continue;
            qu[e] = nu[e];
            qv[e] = nv[e];
            qw[e] = nw[e];
        }
    }
    if (d.T > 0) {
        const int np = d.npml;
        const long qrow = (long)(j - np) * (d.nx - 2 * np);
#pragma unroll
        for (int e = 0; e < NC; ++e) {
            const int c = c0 + e;
            if (c < lo || c >= hi) continue;
            snap_image_w(d, s, qrow, (unsigned)(c - np), dus[e], nu[e]);
        }
    }
}

// Kernel 6: CPML strips with one point per lane, the interior with four.
// Tiles below nt1 are BX x BZR one-point tiles (tile = tj * ntx + ti, only
// those that reach outside the interior are listed); tile nt1 + tb * nx4 + tx
// is the four-column tile of rows npml + tb*BZR .., updating columns
// [x4lo + tx*X4W, x4lo + (tx+1)*X4W) clipped to [npml, nx - npml).
template <int BZR>
__global__ __launch_bounds__(BX * BZR) void k_stepx(const ShotDev* __restrict__ shots,
                                                    const TileRef* __restrict__ tiles,
                                                    const int* __restrict__ base, int rel)
{
    const TileRef tr = tiles[blockIdx.x];
    if (tr.shot < 0) return;
    const ShotDev& d = shots[tr.shot];
    const int s = *base + rel - d.t0;
    if (s < 0 || s >= d.nt) return;
    const int wy = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    if (tr.tile < d.nt1) {
        const int tj = tr.tile / d.ntx;
        const int ti = tr.tile - tj * d.ntx;
        const int j = tj * BZR + wy;
        if (j >= d.nz) return;
        if (d.mode == MODE_REC) rec_one<true>(d, s, j, ti * BX, threadIdx.x);
        else step_one<true>(d, s, j, ti * BX, threadIdx.x);
    } else {
        const int u = tr.tile - d.nt1;
        const int tb = u / d.nx4;
        const int tx = u - tb * d.nx4;
        const int j = d.npml + tb * BZR + wy;
        if (j >= d.nz - d.npml) return;
        const int lo = d.x4lo + tx * X4W;
        step_xn<4>(d, s, j, lo - 4, lo < d.npml ? d.npml : lo, min(lo + X4W, d.x4hi),
                   threadIdx.x);
    }
}

// Kernel 10: kernel 6 with two columns per lane in the interior (124 columns
// per wave, half the registers of four).
constexpr int XC2 = 2 * (BX - 2);   // columns updated per wave, two columns per lane

template <int BZR>
__global__ __launch_bounds__(BX * BZR) void k_stepx2(const ShotDev* __restrict__ shots,
                                                     const TileRef* __restrict__ tiles,
                                                     const int* __restrict__ base, int rel)
{
    const TileRef tr = tiles[blockIdx.x];
    if (tr.shot < 0) return;
    const ShotDev& d = shots[tr.shot];
    const int s = *base + rel - d.t0;
    if (s < 0 || s >= d.nt) return;
    const int wy = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    if (tr.tile < d.nt1) {
        const int tj = tr.tile / d.ntx;
        const int ti = tr.tile - tj * d.ntx;
        const int j = tj * BZR + wy;
        if (j >= d.nz) return;
        if (d.mode == MODE_REC) rec_one<true>(d, s, j, ti * BX, threadIdx.x);
        else step_one<true>(d, s, j, ti * BX, threadIdx.x);
    } else {
        const int u = tr.tile - d.nt1;
        const int tb = u / d.nx4;
        const int tx = u - tb * d.nx4;
        const int j = d.npml + tb * BZR + wy;
        if (j >= d.nz - d.npml) return;
        const int lo = d.x4lo + tx * XC2;
        step_xn<2>(d, s, j, lo - 2, lo < d.npml ? d.npml : lo, min(lo + XC2, d.x4hi),
                   threadIdx.x);
    }
}

// Neighbourhood of a point from an LDS array of row width W.
__device__ __forceinline__ void nbh_lds(const float* a, int W, Nbh& o)
{
    o.nw = a[-W - 1]; o.n = a[-W]; o.ne = a[-W + 1];
    o.w = a[-1];      o.c = a[0];  o.e = a[1];
    o.sw = a[W - 1];  o.s = a[W];  o.se = a[W + 1];
}

// ------------------------------------------- two steps per launch, wave rows
// A block of four wavefronts owns an output tile of 60 columns x OZ rows and
// advances it two time levels in one launch:
//   phase 0  stage time level s of U, Vs, Ws on the tile plus a 4-point halo
//            (OZ+8 rows x 68 columns) in LDS, a wavefront per row;
//   phase 1  level s+1 on the tile plus a 2-point halo (OZ+4 rows x 64
//            columns, lane l <-> column x0-2+l) from LDS into LDS; tile
//            points also go to global memory (the next launch's previous
//            level), with their receivers, snapshots and imaging of step s;
//   phase 2  level s+2 on the tile (lanes 2..61), step s+1's receivers,
//            snapshots and imaging.
// Per output point and two steps this reads level s with its halo, level s-1
// and rx once and writes levels s+1 and s+2 (about 52 B against 80 B for two
// one-step launches), with the neighbourhood reuse in LDS instead of L2, and
// halves the launches.  Time level t lives in f[t % 4], so no block
// overwrites a level another block still reads.  CPML memories have two sets
// (planes 0-8, 9-17): the launch at step s reads set (s/2)%2 and writes the
// other.  Phase 1 steps the memories of every point it computes (halo
// included) in registers and drops them; phase 2 reloads its point's memories,
// repeats the level s+1 update from the staged level s, steps them to level
// s+2 and stores them (a launch that runs one step stores the s+1 set from
// phase 1).
constexpr int X2W = 60;   // output columns of a two-step wave-row tile

// The CPML memories see U without the source injection that the previous step
// folded into it (fwd_PDE_exact_MNAD.m:95-111): remove it from the values at
// and around the source.
__device__ __forceinline__ void pre_injection(const ShotDev& d, int s, int j, int i, float& uc,
                                              float& un, float& us, float& ue, float& uw)
{
    if (d.mode != MODE_FWD || d.src_j < 0) return;
    const float a0 = d.src_amp[s];
    if (j == d.src_j) {
        if (i == d.src_i) uc -= a0;
        if (i - 1 == d.src_i) uw -= a0;
        if (i + 1 == d.src_i) ue -= a0;
    }
    if (i == d.src_i) {
        if (j - 1 == d.src_j) un -= a0;
        if (j + 1 == d.src_j) us -= a0;
    }
}

// CPML term of one point for step s (neighbourhoods u, v, w of level s): the
// memories are read from set `pin` and, if `store` >= 0, written to set
// `store` after the update; `upd` also receives the updated memories.
struct PsiPt {
    bool inx, inz;
    long nsx, rowx, nsz, rowz;
    unsigned bx, bz;
};

__device__ __forceinline__ PsiPt psi_point(const ShotDev& d, int j, int i, int i0, unsigned lane)
{
    PsiPt q;
    const int np = d.npml;
    q.inx = (i < np) || (i >= d.nx - np);
    q.inz = (j < np) || (j >= d.nz - np);
    q.nsx = (long)d.nz * 2 * np;
    q.rowx = (long)j * 2 * np;
    q.bx = 4u * ((i < np) ? (unsigned)i : (unsigned)(i - (d.nx - 2 * np)));
    q.nsz = (long)2 * np * d.pitch;
    q.rowz = (long)((j < np) ? j : j - (d.nz - 2 * np)) * d.pitch + i0;
    q.bz = 4u * lane;
    return q;
}

template <int OZ, int NW>
__global__ __launch_bounds__(BX * NW) void k_step2w(const ShotDev* __restrict__ shots,
                                                    const TileRef* __restrict__ tiles,
                                                    const int* __restrict__ base, int rel)
{
    constexpr int R0 = OZ + 8, R1 = OZ + 4, LW = BX + 4;
    static_assert(R1 % NW == 0, "phase-1 rows must split evenly over the wavefronts");
    __shared__ float s0[3 * R0 * LW];   // level s, rows z0-4 .., columns x0-4 ..
    __shared__ float s1[3 * R1 * BX];   // level s+1, rows z0-2 .., columns x0-2 ..

    const TileRef tr = tiles[blockIdx.x];
    if (tr.shot < 0) return;
    const ShotDev& d = shots[tr.shot];
    const int s = *base + rel - d.t0;
    if (s < 0 || s >= d.nt) return;   // uniform over the block
    const int tj = tr.tile / d.ntx;
    const int ti = tr.tile - tj * d.ntx;
    const int x0 = ti * X2W, z0 = tj * OZ;
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const unsigned lane = threadIdx.x;
    const int nx = d.nx, nz = d.nz, np = d.npml;
    const long P = d.pitch;
    const bool second = s + 1 < d.nt;
    const int pin = (s >> 1) & 1;

    // Per-point global inputs of this wave's phase-1 rows (r1 = wv + 4k), issued
    // before the staging so that their latency overlaps it: rx, level s-1,
    // receiver slot.  Phase 2 reuses rx (same thread, same point).
    constexpr int K1 = R1 / NW;
    float rxk[K1], pu[K1], pv[K1], pw[K1];
    int rk[K1];
#pragma unroll
    for (int k = 0; k < K1; ++k) {
        const int j = z0 - 2 + wv + NW * k, i = x0 - 2 + (int)lane;
        rxk[k] = pu[k] = pv[k] = pw[k] = 0.0f;
        rk[k] = -1;
        if (j >= 0 && j < nz && i >= 0 && i < nx) {
            const long org = (long)(j + MNAD_ZO) * P + (x0 - 2 + MNAD_XO);   // (j, x0-2)
            rxk[k] = (gp(d.rx) + org)[lane];
            pu[k] = (gp(d.f[(s + 3) & 3][0]) + org)[lane];
            pv[k] = (gp(d.f[(s + 3) & 3][1]) + org)[lane];
            pw[k] = (gp(d.f[(s + 3) & 3][2]) + org)[lane];
            if (d.nrec > 0) {
                const int off = __builtin_amdgcn_readfirstlane(d.rrow[j]);
                if (off >= 0) rk[k] = uni(gp(d.rcol) + off + x0 - 2)[lane];
            }
        }
    }

    // ---- phase 0: stage level s (all rows of the wave issued together)
    constexpr int K0 = (R0 + NW - 1) / NW;
#pragma unroll
    for (int k = 0; k < K0; ++k) {
        const int r = wv + NW * k;
        if (r >= R0) break;
        const int j = z0 - 4 + r;
        float a[3] = {0.0f, 0.0f, 0.0f}, e[3] = {0.0f, 0.0f, 0.0f};
        if (j >= 0 && j < nz) {
            const long org = (long)(j + MNAD_ZO) * P + (x0 - 4 + MNAD_XO);
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                const gfloat* src = gp(d.f[s & 3][f]) + org;
                a[f] = src[lane];
                if (lane < 4) e[f] = (src + BX)[lane];
            }
        }
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            s0[(f * R0 + r) * LW + lane] = a[f];
            if (lane < 4) s0[(f * R0 + r) * LW + BX + lane] = e[f];
        }
    }
    __syncthreads();

    // ---- phase 1: level s+1 on the tile and its 2-point halo
#pragma unroll
    for (int k = 0; k < K1; ++k) {
        const int r1 = wv + NW * k;
        const int j = z0 - 2 + r1, i = x0 - 2 + (int)lane;
        float n0 = 0.0f, n1 = 0.0f, n2 = 0.0f;
        if (j >= 0 && j < nz && i >= 0 && i < nx) {
            const float* c = s0 + (r1 + 2) * LW + lane + 2;
            Nbh u, v, w;
            nbh_lds(c, LW, u);
            nbh_lds(c + R0 * LW, LW, v);
            nbh_lds(c + 2 * R0 * LW, LW, w);
            const float vww = c[R0 * LW - 2], vee = c[R0 * LW + 2];
            const float wnn = c[2 * R0 * LW - 2 * LW], wss = c[2 * R0 * LW + 2 * LW];
            const long org = (long)(j + MNAD_ZO) * P + (x0 - 2 + MNAD_XO);   // (j, x0-2)
            const float rx = rxk[k];
            const int r = rk[k];
            const float rz = rx * d.kap;
            float du, dv, dw;
            mnad_incr(d.ic, rx, u, v, vww, vee, w, wnn, wss, du, dv, dw);
            const bool own = j >= z0 && j < z0 + OZ && lane >= 2 && lane < BX - 2;
            const PsiPt q = psi_point(d, j, i, x0 - 2, lane);
            if (q.inx || q.inz) {
                float uc = u.c, un = u.n, us = u.s, ue = u.e, uw = u.w;
                pre_injection(d, s, j, i, uc, un, us, ue, uw);
                const bool keep = own && !second;   // a one-step launch stores the s+1 set here
                if (q.inx) {
                    PsiW m;
                    psi_load(m, gp(d.psx) + pin * 9 * q.nsx + q.rowx, gp(d.cfx) + q.rowx, q.nsx, q.bx);
                    psi_step(m, rx, uw, uc, ue, v.w, v.c, v.e, w.w, w.c, w.e, du, dv, dw);
                    if (keep) psi_store(m, gp(d.psx) + (pin ^ 1) * 9 * q.nsx + q.rowx, q.nsx, q.bx);
                }
                if (q.inz) {
                    PsiW m;
                    psi_load(m, gp(d.psz) + pin * 9 * q.nsz + q.rowz, gp(d.cfz) + q.rowz, q.nsz, q.bz);
                    psi_step(m, rz, un, uc, us, v.n, v.c, v.s, w.n, w.c, w.s, du, dv, dw);
                    if (keep) psi_store(m, gp(d.psz) + (pin ^ 1) * 9 * q.nsz + q.rowz, q.nsz, q.bz);
                }
            }
            n0 = (2.0f * u.c + du) - pu[k];
            n1 = (2.0f * v.c + dv) - pv[k];
            n2 = (2.0f * w.c + dw) - pw[k];
            if (d.mode == MODE_FWD) {
                if (j == d.src_j && i == d.src_i) n0 += d.src_amp[s + 1];
                if (own && r >= 0) (gp(d.gather) + (long)s * d.nrec)[(unsigned)r] = u.c;
            } else if (r >= 0) {
                n0 -= (gp(d.radj) + (long)s * d.nrec)[(unsigned)r];
            }
            if (own) {
                (gp(d.f[(s + 1) & 3][0]) + org)[lane] = n0;
                (gp(d.f[(s + 1) & 3][1]) + org)[lane] = n1;
                (gp(d.f[(s + 1) & 3][2]) + org)[lane] = n2;
                if (d.T > 0 && !q.inx && !q.inz)
                    snap_image_w(d, s, (long)(j - np) * (nx - 2 * np), (unsigned)(i - np), du, n0);
            }
        }
        s1[(0 * R1 + r1) * BX + lane] = n0;
        s1[(1 * R1 + r1) * BX + lane] = n1;
        s1[(2 * R1 + r1) * BX + lane] = n2;
    }
    __syncthreads();
    if (!second) return;

    // ---- phase 2: level s+2 on the tile; rows with (row - z0 + 2) % 4 == wave,
    // the rows this wave computed in phase 1
#pragma unroll
    for (int k = 0; k < K1; ++k) {
        const int q2 = wv + NW * k - 2;   // this wave's phase-1 row, as a tile row
        if (q2 < 0 || q2 >= OZ) continue;
        const int j = z0 + q2, i = x0 - 2 + (int)lane;
        if (j >= nz) break;
        if (lane < 2 || lane >= BX - 2 || i >= nx) continue;
        const float* c = s1 + (q2 + 2) * BX + lane;
        Nbh u, v, w;
        nbh_lds(c, BX, u);
        nbh_lds(c + R1 * BX, BX, v);
        nbh_lds(c + 2 * R1 * BX, BX, w);
        const float vww = c[R1 * BX - 2], vee = c[R1 * BX + 2];
        const float wnn = c[2 * R1 * BX - 2 * BX], wss = c[2 * R1 * BX + 2 * BX];
        const float* c0 = s0 + (q2 + 4) * LW + lane + 2;   // level s at (j, i)
        const long org = (long)(j + MNAD_ZO) * P + (x0 - 2 + MNAD_XO);
        const float rx = rxk[k];
        const int r = rk[k];
        const float rz = rx * d.kap;
        float du, dv, dw;
        mnad_incr(d.ic, rx, u, v, vww, vee, w, wnn, wss, du, dv, dw);
        const PsiPt q = psi_point(d, j, i, x0 - 2, lane);
        if (q.inx || q.inz) {
            // level s neighbourhood along the axes (staged), pre-injection at step s
            const float* cv = c0 + R0 * LW;
            const float* cw = c0 + 2 * R0 * LW;
            float a_c = c0[0], a_n = c0[-LW], a_s = c0[LW], a_e = c0[1], a_w = c0[-1];
            pre_injection(d, s, j, i, a_c, a_n, a_s, a_e, a_w);
            float uc = u.c, un = u.n, us = u.s, ue = u.e, uw = u.w;
            pre_injection(d, s + 1, j, i, uc, un, us, ue, uw);
            float t0, t1, t2;   // discarded increments of the repeated level s+1 update
            if (q.inx) {
                PsiW m;
                psi_load(m, gp(d.psx) + pin * 9 * q.nsx + q.rowx, gp(d.cfx) + q.rowx, q.nsx, q.bx);
                const float rx0 = rx;
                t0 = t1 = t2 = 0.0f;
                psi_step(m, rx0, a_w, a_c, a_e, cv[-1], cv[0], cv[1], cw[-1], cw[0], cw[1], t0, t1, t2);
                psi_step(m, rx, uw, uc, ue, v.w, v.c, v.e, w.w, w.c, w.e, du, dv, dw);
                psi_store(m, gp(d.psx) + (pin ^ 1) * 9 * q.nsx + q.rowx, q.nsx, q.bx);
            }
            if (q.inz) {
                PsiW m;
                psi_load(m, gp(d.psz) + pin * 9 * q.nsz + q.rowz, gp(d.cfz) + q.rowz, q.nsz, q.bz);
                t0 = t1 = t2 = 0.0f;
                psi_step(m, rz, a_n, a_c, a_s, cv[-LW], cv[0], cv[LW], cw[-LW], cw[0], cw[LW], t0, t1, t2);
                psi_step(m, rz, un, uc, us, v.n, v.c, v.s, w.n, w.c, w.s, du, dv, dw);
                psi_store(m, gp(d.psz) + (pin ^ 1) * 9 * q.nsz + q.rowz, q.nsz, q.bz);
            }
        }
        float n0 = (2.0f * u.c + du) - c0[0];
        const float n1 = (2.0f * v.c + dv) - c0[R0 * LW];
        const float n2 = (2.0f * w.c + dw) - c0[2 * R0 * LW];
        if (d.mode == MODE_FWD) {
            if (j == d.src_j && i == d.src_i) n0 += d.src_amp[s + 2];
            if (r >= 0) (gp(d.gather) + (long)(s + 1) * d.nrec)[(unsigned)r] = u.c;
        } else if (r >= 0) {
            n0 -= (gp(d.radj) + (long)(s + 1) * d.nrec)[(unsigned)r];
        }
        (gp(d.f[(s + 2) & 3][0]) + org)[lane] = n0;
        (gp(d.f[(s + 2) & 3][1]) + org)[lane] = n1;
        (gp(d.f[(s + 2) & 3][2]) + org)[lane] = n2;
        if (d.T > 0 && !q.inx && !q.inz)
            snap_image_w(d, s + 1, (long)(j - np) * (nx - 2 * np), (unsigned)(i - np), du, n0);
    }
}

// ------------------------------------ snapshot / imaging work of a step
// Snapshot / imaging work of step s (snap_image_w), decided once per pass
// instead of once per point: kind 0 none, 1 forward d2u (illumination, and the
// snapshot when the shot keeps them), 2 adjoint imaging, 3 reconstruction
// hand-off slot.  `plane` is the snapshot (or slot) plane of step s.
struct SnapStep {
    int kind;
    void* plane;
};

__device__ __forceinline__ SnapStep snap_step(const ShotDev& d, int s)
{
    SnapStep o{0, nullptr};
    if (d.T <= 0) return o;
    const long ni = (long)(d.nz - 2 * d.npml) * (d.nx - 2 * d.npml);
    const bool f32 = d.sfmt == SNAP_F32 || d.sfmt == SNAP_SLOT;
    if (d.mode == MODE_REC) {
        if ((d.ntf - s) % d.T == 0) o = {3, d.snaps};
    } else if (d.mode == MODE_FWD) {
        if ((s + 1) % d.T == 0) {
            const long k0 = (long)((s + 1) / d.T - 1) * ni;
            o.kind = 1;
            if (d.snaps)
                o.plane = f32 ? (void*)((float*)d.snaps + k0) : (void*)((unsigned short*)d.snaps + k0);
        }
    } else {
        const int tt = d.nt - s;
        if (tt % d.T == 0) {
            const int col = d.shift + tt / d.T;
            if (col <= d.K) {
                const long k0 = d.sfmt == SNAP_SLOT ? 0L : (long)(col - 1) * ni;
                o = {2, f32 ? (void*)((float*)d.snaps + k0) : (void*)((unsigned short*)d.snaps + k0)};
            }
        }
    }
    return o;
}

// ------------------------------------ two steps per launch, register queues (kernel 13)
// Each wavefront owns XQ = 56 output columns: lane l holds column x0 - 4 + l,
// so level s is known on 64 columns, level s+1 on the 60 of lanes 2..61 and
// level s+2 on the 56 of lanes 4..59.  x neighbours come from the adjacent
// lanes (DPP wave shifts), z neighbours from per-lane register queues: the
// wavefront marches down its segment of rows [zA, zB) and at march step t
//   - has level s row r = zA - 4 + t (loaded QD steps ahead),
//   - updates level s+1 on row a = r - 2 (level s rows a-2 .. a+2),
//   - updates level s+2 on row b = r - 4 (level s+1 rows b-2 .. b+2, level s row b).
// Every queue (level s, level s+1, rx, level s-1) is a ring of QP = 5 + QD
// rows in which row j sits in slot (j - zA + 4) mod QP, and the march is
// unrolled by QP, so that every slot is a fixed register (a ring holds only
// the rows still live, so the registers follow the live rows, not QP).  No
// LDS and no barriers: the waves of a workgroup are independent.  Tiles whose
// box reaches within max(npml + 2, 4) cells of the grid edge take the checked
// path with the CPML code (both CPML memory sets as kernel 7: phase 1 reads set
// pin everywhere, phase 2 re-runs the level-s memory step of its own points
// and writes set pin ^ 1); the others are interior tiles whose every load is in
// the grid and whose level s+1 halo points lie outside the CPML, and which run
// without bounds or CPML code (it runs one step body with the rings rotated
// by one slot after each step).  Injections, receivers, snapshots/imaging and
// the point arithmetic are kernel 7's.  Time level t lives in f[t % 4].
constexpr int XQ = BX - 8;
// QD: prefetch distance, in march steps; QP = 5 + QD: ring length.  The
// launch runs QD = 2 on both paths: at the C5 grid a distance of 3 or 4 (110
// and 120 VGPRs, the same 4 waves per SIMD) measured 0-1 % slower.
template <int QD>
__host__ __device__ constexpr int qslot(int ph, int off)
{
    constexpr int QP = 5 + QD;
    return ((ph + off) % QP + QP) % QP;
}

// Buffer access to the field planes: a shot's 3 nlev planes are one
// allocation (plane 3 level + field at byte (3 level + field) * plane bytes),
// so one resource in scalar registers covers all of them and rx has its own.
// An access gives the level, row and wave as a uniform byte offset (soffset)
// and the field and lane as a per-lane constant (voffset vo[f] = 4 lane + f
// plane bytes), with no per-lane address arithmetic.  The values are read
// from the descriptor once, before the march stores anything: the compiler
// cannot tell that the buffer stores leave the descriptor alone, and would
// otherwise read it again after each of them.
struct QCtx {
    __amdgpu_buffer_rsrc_t rs, rr;   // field planes, rx
    unsigned pitch4;                 // row bytes
    unsigned b0, bm, b1, b2, bx;     // (0, x0 - 4) of levels s, s-1, s+1, s+2 and of rx, in bytes
    unsigned vo[3];                  // per lane: 4 lane + field * plane bytes
};

__device__ __forceinline__ __amdgpu_buffer_rsrc_t qres(const void* p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ float qld(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned vo)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, (int)off, 0));
}

__device__ __forceinline__ void qst(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned vo, float x)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, (int)vo, (int)off, 0);
}

// A lane offset that the buffer range check rejects: a store through it is
// dropped.  The interior path masks its stores this way instead of branching
// around them (every state allocation is under 2 GiB, the run_batch check).
constexpr unsigned QOFF = 0x80000000u;

// Receiver-row table entry j, read through the constant address space: the
// table is never written during a launch, so the read is a scalar load,
// which counts on lgkmcnt and does not make the wave wait for its in-flight
// prefetches (vector loads complete in order, so waiting for a vector load
// of the table would wait for all of them).
__device__ __forceinline__ int q_rrow(const ShotDev& d, int j)
{
    return ((const __attribute__((address_space(4))) int*)d.rrow)[j];
}

template <int QD>
struct March {
    static constexpr int QP = 5 + QD;
    float q0[3][QP];   // level s (U, Vs, Ws)
    float q1[3][QP];   // level s+1
    float qr[QP];      // rx
    float qp[3][QP];   // level s-1
};

// Level s+1 at row a (phase 1), step t with PH = t mod QP: level s row a + k
// in slot qslot<QD>(PH, k - 2); rx, level s-1 and the new level s+1 of row a in
// slot qslot<QD>(PH, -2).  PURE: no receiver rows in reach (the interior main loop).
template <int QD, bool EDGE, bool K1, int PH, bool PURE = false>
__device__ __forceinline__ void q_phase1(const ShotDev& d, March<QD>& m, int s, int a, int zA, int zB,
                                         int i, unsigned lane, const QCtx& qc, bool colin, bool ownc,
                                         bool xedge, float amp1, bool second)
{
    constexpr int sM2 = qslot<QD>(PH, -4), sM1 = qslot<QD>(PH, -3), sC = qslot<QD>(PH, -2), sP1 = qslot<QD>(PH, -1),
                  sP2 = qslot<QD>(PH, 0);
    const int nz = d.nz, nx = d.nx, np = d.npml;
    const bool rec = d.mode == MODE_REC, adj = d.mode == MODE_ADJ;
    Nbh u, v, w;
    u.n = m.q0[0][sM1]; u.c = m.q0[0][sC]; u.s = m.q0[0][sP1];
    v.n = m.q0[1][sM1]; v.c = m.q0[1][sC]; v.s = m.q0[1][sP1];
    w.n = m.q0[2][sM1]; w.c = m.q0[2][sC]; w.s = m.q0[2][sP1];
    u.nw = lane_prev(u.n); u.ne = lane_next(u.n); u.w = lane_prev(u.c); u.e = lane_next(u.c);
    u.sw = lane_prev(u.s); u.se = lane_next(u.s);
    v.nw = lane_prev(v.n); v.ne = lane_next(v.n); v.w = lane_prev(v.c); v.e = lane_next(v.c);
    v.sw = lane_prev(v.s); v.se = lane_next(v.s);
    w.nw = lane_prev(w.n); w.ne = lane_next(w.n); w.w = lane_prev(w.c); w.e = lane_next(w.c);
    w.sw = lane_prev(w.s); w.se = lane_next(w.s);
    const float vww = lane_prev(v.w), vee = lane_next(v.e);
    const float wnn = m.q0[2][sM2], wss = m.q0[2][sP2];
    const float rx = m.qr[sC];
    float du, dv, dw;
    if constexpr (K1) {   // dx == dz: the weights as literals
        constexpr IncrC c1 = incr_consts_of(1.0);
        mnad_incr(c1, rx, u, v, vww, vee, w, wnn, wss, du, dv, dw);
    } else {
        mnad_incr(d.ic, rx, u, v, vww, vee, w, wnn, wss, du, dv, dw);
    }
    const bool rowin = !EDGE || (a >= 0 && a < nz);
    const bool own = ownc && a >= zA && a < zB;
    if (EDGE && !rec) {   // a reconstruction runs inside R, which holds no CPML memories
        const bool zedge = a < np || a >= nz - np;
        if ((xedge || zedge) && rowin) {
            const PsiPt q = psi_point(d, a, i, i - (int)lane, lane);
            if (colin && (q.inx || q.inz)) {
                const int pin = (s >> 1) & 1;
                const float rz = rx * d.kap;
                float a_c = u.c, a_n = u.n, a_s = u.s, a_e = u.e, a_w = u.w;
                pre_injection(d, s, a, i, a_c, a_n, a_s, a_e, a_w);
                const bool keep = own && !second;   // a one-step launch stores the s+1 set
                if (q.inx) {
                    PsiW mm;
                    psi_load(mm, gp(d.psx) + pin * 9 * q.nsx + q.rowx, gp(d.cfx) + q.rowx, q.nsx, q.bx);
                    psi_step(mm, rx, a_w, a_c, a_e, v.w, v.c, v.e, w.w, w.c, w.e, du, dv, dw);
                    if (keep) psi_store(mm, gp(d.psx) + (pin ^ 1) * 9 * q.nsx + q.rowx, q.nsx, q.bx);
                }
                if (q.inz) {
                    PsiW mm;
                    psi_load(mm, gp(d.psz) + pin * 9 * q.nsz + q.rowz, gp(d.cfz) + q.rowz, q.nsz, q.bz);
                    psi_step(mm, rz, a_n, a_c, a_s, v.n, v.c, v.s, w.n, w.c, w.s, du, dv, dw);
                    if (keep) psi_store(mm, gp(d.psz) + (pin ^ 1) * 9 * q.nsz + q.rowz, q.nsz, q.bz);
                }
            }
        }
    }
    float n0 = (2.0f * u.c + du) - m.qp[0][sC];
    float n1 = (2.0f * v.c + dv) - m.qp[1][sC];
    float n2 = (2.0f * w.c + dw) - m.qp[2][sC];
    if (!adj && a == d.src_j && i == d.src_i) n0 += amp1;
    if (!PURE && adj && d.nrec > 0 && rowin) {   // the adjoint's receiver injection
        const int off = q_rrow(d, a);
        if (off >= 0 && colin) {
            const int r = uni(gp(d.rcol) + off)[(unsigned)i];
            if (r >= 0) n0 -= (gp(d.radj) + (long)s * d.nrec)[(unsigned)r];
        }
    }
    bool put = own;
    if (EDGE) {
        if (!(rowin && colin)) {
            n0 = n1 = n2 = 0.0f;
        } else if (rec) {
            const bool inr = in_rec_region(a, i, nz, nx, np);
            put = put && inr;
            if (ring_index(a, i, nz, nx, np) >= 0) {   // decoded before the launch
                const unsigned o = qc.b1 + (unsigned)a * qc.pitch4;
                n0 = qld(qc.rs, o, qc.vo[0]);
                n1 = qld(qc.rs, o, qc.vo[1]);
                n2 = qld(qc.rs, o, qc.vo[2]);
            } else if (!inr) {
                n0 = n1 = n2 = 0.0f;   // R never reads it
            }
        }
        if (put) {
            const unsigned o = qc.b1 + (unsigned)a * qc.pitch4;
            qst(qc.rs, o, qc.vo[0], n0);
            qst(qc.rs, o, qc.vo[1], n1);
            qst(qc.rs, o, qc.vo[2], n2);
        }
    } else {   // no branch: the stores of lanes and rows that are not this wave's are dropped
        const unsigned o = qc.b1 + (unsigned)a * qc.pitch4, msk = put ? 0u : QOFF;
        qst(qc.rs, o, qc.vo[0] + msk, n0);
        qst(qc.rs, o, qc.vo[1] + msk, n1);
        qst(qc.rs, o, qc.vo[2] + msk, n2);
    }
    m.q1[0][sC] = n0;
    m.q1[1][sC] = n1;
    m.q1[2][sC] = n2;
}

// Level s+2 at row b = a - 2 (phase 2), step t with PH = t mod QP: level s+1
// row b + k in slot qslot<QD>(PH, k - 4), rx and level s of row b in slot
// qslot<QD>(PH, -4).  bm0..bm2: level s row b - 1 (its slot already holds the
// prefetched row; the edge path's repeated CPML step reads it).
template <int QD, bool EDGE, bool K1, int PH, bool PURE = false>
__device__ __forceinline__ void q_phase2(const ShotDev& d, March<QD>& m, int s, int b, int i,
                                         unsigned lane, const QCtx& qc, bool ownc, bool xedge,
                                         float amp2, float bm0, float bm1, float bm2)
{
    constexpr int tM2 = qslot<QD>(PH, -6), tM1 = qslot<QD>(PH, -5), tC = qslot<QD>(PH, -4), tP1 = qslot<QD>(PH, -3),
                  tP2 = qslot<QD>(PH, -2);
    const int nz = d.nz, nx = d.nx, np = d.npml;
    const bool rec = d.mode == MODE_REC, adj = d.mode == MODE_ADJ;
    Nbh u, v, w;
    u.n = m.q1[0][tM1]; u.c = m.q1[0][tC]; u.s = m.q1[0][tP1];
    v.n = m.q1[1][tM1]; v.c = m.q1[1][tC]; v.s = m.q1[1][tP1];
    w.n = m.q1[2][tM1]; w.c = m.q1[2][tC]; w.s = m.q1[2][tP1];
    u.nw = lane_prev(u.n); u.ne = lane_next(u.n); u.w = lane_prev(u.c); u.e = lane_next(u.c);
    u.sw = lane_prev(u.s); u.se = lane_next(u.s);
    v.nw = lane_prev(v.n); v.ne = lane_next(v.n); v.w = lane_prev(v.c); v.e = lane_next(v.c);
    v.sw = lane_prev(v.s); v.se = lane_next(v.s);
    w.nw = lane_prev(w.n); w.ne = lane_next(w.n); w.w = lane_prev(w.c); w.e = lane_next(w.c);
    w.sw = lane_prev(w.s); w.se = lane_next(w.s);
    const float vww = lane_prev(v.w), vee = lane_next(v.e);
    const float wnn = m.q1[2][tM2], wss = m.q1[2][tP2];
    const float rx = m.qr[tC];
    const float c0 = m.q0[0][tC], cv = m.q0[1][tC], cw = m.q0[2][tC];   // level s at (b, i)
    float du, dv, dw;
    if constexpr (K1) {   // dx == dz: the weights as literals
        constexpr IncrC c1 = incr_consts_of(1.0);
        mnad_incr(c1, rx, u, v, vww, vee, w, wnn, wss, du, dv, dw);
    } else {
        mnad_incr(d.ic, rx, u, v, vww, vee, w, wnn, wss, du, dv, dw);
    }
    if (EDGE && !rec) {
        const bool zedge = b < np || b >= nz - np;
        if (xedge || zedge) {
            // level s along the axes at (b, i), x neighbours while every lane is active
            const float c0e = lane_next(c0), c0w = lane_prev(c0), cve = lane_next(cv),
                        cvw = lane_prev(cv), cwe = lane_next(cw), cww = lane_prev(cw);
            const PsiPt q = psi_point(d, b, i, i - (int)lane, lane);
            if (ownc && (q.inx || q.inz)) {
                const int pin = (s >> 1) & 1;
                const float rz = rx * d.kap;
                float a_c = c0, a_n = bm0, a_s = m.q0[0][tP1], a_e = c0e, a_w = c0w;
                pre_injection(d, s, b, i, a_c, a_n, a_s, a_e, a_w);
                float uc1 = u.c, un1 = u.n, us1 = u.s, ue1 = u.e, uw1 = u.w;
                pre_injection(d, s + 1, b, i, uc1, un1, us1, ue1, uw1);
                float t0, t1, t2;   // discarded increments of the repeated level s+1 update
                if (q.inx) {
                    PsiW mm;
                    psi_load(mm, gp(d.psx) + pin * 9 * q.nsx + q.rowx, gp(d.cfx) + q.rowx, q.nsx, q.bx);
                    t0 = t1 = t2 = 0.0f;
                    psi_step(mm, rx, a_w, a_c, a_e, cvw, cv, cve, cww, cw, cwe, t0, t1, t2);
                    psi_step(mm, rx, uw1, uc1, ue1, v.w, v.c, v.e, w.w, w.c, w.e, du, dv, dw);
                    psi_store(mm, gp(d.psx) + (pin ^ 1) * 9 * q.nsx + q.rowx, q.nsx, q.bx);
                }
                if (q.inz) {
                    PsiW mm;
                    psi_load(mm, gp(d.psz) + pin * 9 * q.nsz + q.rowz, gp(d.cfz) + q.rowz, q.nsz, q.bz);
                    t0 = t1 = t2 = 0.0f;
                    psi_step(mm, rz, a_n, a_c, a_s, bm1, cv, m.q0[1][tP1], bm2, cw, m.q0[2][tP1],
                             t0, t1, t2);
                    psi_step(mm, rz, un1, uc1, us1, v.n, v.c, v.s, w.n, w.c, w.s, du, dv, dw);
                    psi_store(mm, gp(d.psz) + (pin ^ 1) * 9 * q.nsz + q.rowz, q.nsz, q.bz);
                }
            }
        }
    }
    float n0 = (2.0f * u.c + du) - c0;
    const float n1 = (2.0f * v.c + dv) - cv;
    const float n2 = (2.0f * w.c + dw) - cw;
    if (!adj && b == d.src_j && i == d.src_i) n0 += amp2;
    if (!PURE && adj && d.nrec > 0) {   // the adjoint's receiver injection
        const int off = q_rrow(d, b);
        if (off >= 0 && ownc) {
            const int r = uni(gp(d.rcol) + off)[(unsigned)i];
            if (r >= 0) n0 -= (gp(d.radj) + (long)(s + 1) * d.nrec)[(unsigned)r];
        }
    }
    if (EDGE) {
        // a reconstruction writes R only (its ring cells were decoded before the launch)
        if (ownc && (!rec || in_rec_region(b, i, nz, nx, np))) {
            const unsigned o = qc.b2 + (unsigned)b * qc.pitch4;
            qst(qc.rs, o, qc.vo[0], n0);
            qst(qc.rs, o, qc.vo[1], n1);
            qst(qc.rs, o, qc.vo[2], n2);
        }
    } else {   // no branch: the stores of lanes that are not this wave's are dropped
        const unsigned o = qc.b2 + (unsigned)b * qc.pitch4, msk = ownc ? 0u : QOFF;
        qst(qc.rs, o, qc.vo[0] + msk, n0);
        qst(qc.rs, o, qc.vo[1] + msk, n1);
        qst(qc.rs, o, qc.vo[2] + msk, n2);
    }
}

// Level s row j into slot sl (zeros off the grid on the edge path).
// msk: QOFF on the interior path when the row is not needed (the load is
// dropped by the range check and reads 0).
template <int QD, bool EDGE>
__device__ __forceinline__ void q_load0(const ShotDev& d, March<QD>& m, int s, int j, int sl,
                                        const QCtx& qc, unsigned msk = 0u)
{
    if (!EDGE || (j >= 0 && j < d.nz)) {
        const unsigned o = qc.b0 + (unsigned)j * qc.pitch4;
        m.q0[0][sl] = qld(qc.rs, o, qc.vo[0] + msk);
        m.q0[1][sl] = qld(qc.rs, o, qc.vo[1] + msk);
        m.q0[2][sl] = qld(qc.rs, o, qc.vo[2] + msk);
    } else {
        m.q0[0][sl] = m.q0[1][sl] = m.q0[2][sl] = 0.0f;
    }
}

// Loads of march step t (PH = t mod QP): level s row r + QD, and rx and level
// s-1 of row r - 2 + QD (the level-s+1 row of step t + QD).
template <int QD, bool EDGE, int PH>
__device__ __forceinline__ void q_loads(const ShotDev& d, March<QD>& m, int s, int r, int t, int NR,
                                        const QCtx& qc, bool colin)
{
    const int a = r - 2 + QD;
    constexpr int sl = qslot<QD>(PH, QD - 2);
    if constexpr (!EDGE) {
        // The interior path issues every load of the march without a branch
        // (a branch around a load leaves the wait counts unknown after it).
        // A load whose row the march does not use -- level s past row zB + 3
        // in the last QD steps, rx and level s-1 outside rows zA - 2 .. zB + 1
        // in the first 4 - QD and the last QD steps -- goes to the offset the
        // range check drops, so it costs no memory traffic.
        const bool want = t + QD < NR;
        q_load0<QD, false>(d, m, s, r + QD, qslot<QD>(PH, QD), qc, want ? 0u : QOFF);
        const unsigned o = (unsigned)a * qc.pitch4, mk = (want && t + QD >= 4) ? 0u : QOFF;
        m.qr[sl] = qld(qc.rr, qc.bx + o, qc.vo[0] + mk);
        m.qp[0][sl] = qld(qc.rs, qc.bm + o, qc.vo[0] + mk);
        m.qp[1][sl] = qld(qc.rs, qc.bm + o, qc.vo[1] + mk);
        m.qp[2][sl] = qld(qc.rs, qc.bm + o, qc.vo[2] + mk);
    } else if (t + QD < NR) {
        q_load0<QD, EDGE>(d, m, s, r + QD, qslot<QD>(PH, QD), qc);
        if (t + QD >= 4) {
            if (a >= 0 && a < d.nz && colin) {
                const unsigned o = (unsigned)a * qc.pitch4;
                m.qr[sl] = qld(qc.rr, qc.bx + o, qc.vo[0]);
                m.qp[0][sl] = qld(qc.rs, qc.bm + o, qc.vo[0]);
                m.qp[1][sl] = qld(qc.rs, qc.bm + o, qc.vo[1]);
                m.qp[2][sl] = qld(qc.rs, qc.bm + o, qc.vo[2]);
            } else {
                m.qr[sl] = m.qp[0][sl] = m.qp[1][sl] = m.qp[2][sl] = 0.0f;
            }
        }
    }
}

// March step t with the slots of phase PH = t mod QP.
template <int QD, bool EDGE, bool K1, int PH>
__device__ __forceinline__ void q_step(const ShotDev& d, March<QD>& m, int s, int t, int NR, int zA,
                                       int zB, int i, unsigned lane, const QCtx& qc, bool colin,
                                       bool ownc, bool xedge, float amp1, float amp2, bool second)
{
    const int r = zA - 4 + t;
    // level s row r - 5 = b - 1, whose slot the load below refills (edge path only)
    constexpr int sb = qslot<QD>(PH, QD);
    const float bm0 = EDGE ? m.q0[0][sb] : 0.0f, bm1 = EDGE ? m.q0[1][sb] : 0.0f,
                bm2 = EDGE ? m.q0[2][sb] : 0.0f;
    q_loads<QD, EDGE, PH>(d, m, s, r, t, NR, qc, colin);
    if (t >= 4)
        q_phase1<QD, EDGE, K1, PH>(d, m, s, r - 2, zA, zB, i, lane, qc, colin, ownc, xedge, amp1, second);
    if (t >= 8 && second)
        q_phase2<QD, EDGE, K1, PH>(d, m, s, r - 4, i, lane, qc, ownc, xedge, amp2, bm0, bm1, bm2);
}

// A main-loop step of the interior path (t >= 8, both phases, no adjoint
// receiver rows near the segment): loads, phase 1 and phase 2 with no branch,
// so that the wait counts stay exact and a step waits only for the loads
// issued QD steps earlier.
template <int QD, bool K1, int PH>
__device__ __forceinline__ void q_step_main(const ShotDev& d, March<QD>& m, int s, int t, int NR,
                                            int zA, int zB, int i, unsigned lane, const QCtx& qc,
                                            bool ownc, float amp1, float amp2)
{
    const int r = zA - 4 + t;
    q_loads<QD, false, PH>(d, m, s, r, t, NR, qc, true);
    q_phase1<QD, false, K1, PH, true>(d, m, s, r - 2, zA, zB, i, lane, qc, true, ownc, false, amp1, true);
    q_phase2<QD, false, K1, PH, true>(d, m, s, r - 4, i, lane, qc, ownc, false, amp2, 0.0f, 0.0f, 0.0f);
}

// Steps T = 0 .. 7 of the march, each with the slots of phase T mod QP.
template <int QD, bool EDGE, bool K1, int... T>
__device__ __forceinline__ void q_prologue(const ShotDev& d, March<QD>& m, int s, int NR, int zA,
                                           int zB, int i, unsigned lane, const QCtx& qc, bool colin,
                                           bool ownc, bool xedge, float amp1, float amp2,
                                           bool second, std::integer_sequence<int, T...>)
{
    (q_step<QD, EDGE, K1, T % (5 + QD)>(d, m, s, T, NR, zA, zB, i, lane, qc, colin, ownc, xedge,
                                        amp1, amp2, second),
     ...);
}

// One round of the main loop: steps t + K (K = 0 .. QP-1, t = 8 + n QP), each
// with the slots of phase (8 + K) mod QP, up to step NR - 1.  Returns false
// once it reached NR.
template <int QD, bool K1, int... K>
__device__ __forceinline__ bool q_main_round(const ShotDev& d, March<QD>& m, int s, int t, int NR,
                                             int zA, int zB, int i, unsigned lane, const QCtx& qc,
                                             bool ownc, float amp1, float amp2,
                                             std::integer_sequence<int, K...>)
{
    return ((t + K < NR ? (q_step_main<QD, K1, (8 + K) % (5 + QD)>(d, m, s, t + K, NR, zA, zB, i,
                                                                   lane, qc, ownc, amp1, amp2),
                           true)
                        : false) &&
            ...);
}

template <int QD, bool EDGE, bool K1>
__device__ __forceinline__ void q_march(const ShotDev& d, int s, int x0, int zA, int zB, unsigned lane)
{
    static_assert(QD >= 2 && QD <= 4, "the march's first loads assume a prefetch distance of 2 to 4");
    constexpr int QP = 5 + QD;
    const int nx = d.nx, np = d.npml;
    const int i = x0 - 4 + (int)lane;
    const bool colin = !EDGE || (i >= 0 && i < nx);
    const bool ownc = lane >= 4 && lane < BX - 4 && (!EDGE || i < nx);
    const bool xedge = EDGE && (x0 - 4 < np || x0 + XQ + 4 > nx - np);
    const bool second = s + 1 < d.nt;
    const bool rec = d.mode == MODE_REC;
    // injections: the forward's a_src w into level s+1 (src_amp[s + 1]); a
    // reconstruction takes the forward's injection into level c+1 back out
    // of level c-1 (c = ntf - 1 - s: src_amp[c + 1])
    float amp1 = 0.0f, amp2 = 0.0f;
    if (d.src_j >= 0 && d.mode == MODE_FWD) {
        amp1 = d.src_amp[s + 1];
        amp2 = second ? d.src_amp[s + 2] : 0.0f;
    } else if (d.src_j >= 0 && rec) {
        amp1 = d.src_amp[d.ntf - s];
        amp2 = second ? d.src_amp[d.ntf - 1 - s] : 0.0f;
    }
    // plane of level s + k: forward and adjoint (s + k) mod 4, reconstruction
    // (ntf - 1 - s - k) mod 4 (the forward's levels backwards)
    auto plane = [&](int k) { return rec ? (d.ntf - 1 - s - k) & 3 : (s + k) & 3; };
    QCtx qc;
    {
        const unsigned pb = d.pbytes;   // the host checks the planes are one allocation
        const unsigned xo4 = 4u * ((unsigned)MNAD_ZO * (unsigned)d.pitch + (unsigned)(x0 - 4 + MNAD_XO));
        qc.pitch4 = 4u * (unsigned)d.pitch;
        qc.rs = qres(d.f[0][0], 3u * (unsigned)d.nlev * pb);
        qc.rr = qres(d.rx, pb);
        qc.b0 = 3u * (unsigned)plane(0) * pb + xo4;
        qc.bm = 3u * (unsigned)plane(-1) * pb + xo4;
        qc.b1 = 3u * (unsigned)plane(1) * pb + xo4;
        qc.b2 = 3u * (unsigned)plane(2) * pb + xo4;
        qc.bx = xo4;
        qc.vo[0] = 4u * lane;
        qc.vo[1] = 4u * lane + pb;
        qc.vo[2] = 4u * lane + 2u * pb;
    }
    const int NR = zB - zA + 8;   // march steps: level s rows zA-4 .. zB+3
    March<QD> m;
    // level s rows zA-4 .. zA-4+QD-1 (the first QD steps' rows)
#pragma unroll
    for (int k = 0; k < QD; ++k) q_load0<QD, EDGE>(d, m, s, zA - 4 + k, k, qc);
    // adjoint receiver rows among the rows the march updates (zA-2 .. zB+1)
    const bool rcv = d.mode == MODE_ADJ && d.nrec > 0 && d.rlo <= zB + 1 && d.rhi >= zA - 2;
    if (!EDGE && second && !rcv) {
        // unrolled by QP: the slots are fixed registers.  Steps 0 .. 7 (loads,
        // then phase 1 from step 4), then the branch-free main loop in rounds
        // of QP steps.
        q_prologue<QD, EDGE, K1>(d, m, s, NR, zA, zB, i, lane, qc, colin, ownc, xedge, amp1, amp2,
                                 second, std::make_integer_sequence<int, 8>{});
        for (int t = 8; t < NR; t += QP)   // NR = seg + 8 >= 16
            if (!q_main_round<QD, K1>(d, m, s, t, NR, zA, zB, i, lane, qc, ownc, amp1, amp2,
                                      std::make_integer_sequence<int, QP>{}))
                break;
    } else {
        // One step body (PH = 0) with the rings rotated by one slot after each
        // step: the edge path, and interior blocks with adjoint receiver rows
        // or a last launch of one step.
        for (int t = 0; t < NR; ++t) {
            q_step<QD, EDGE, K1, 0>(d, m, s, t, NR, zA, zB, i, lane, qc, colin, ownc, xedge, amp1,
                                    amp2, second);
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                const float a0 = m.q0[f][0], b0 = m.q1[f][0], p0 = m.qp[f][0];
#pragma unroll
                for (int k = 0; k < QP - 1; ++k) {
                    m.q0[f][k] = m.q0[f][k + 1];
                    m.q1[f][k] = m.q1[f][k + 1];
                    m.qp[f][k] = m.qp[f][k + 1];
                }
                m.q0[f][QP - 1] = a0;
                m.q1[f][QP - 1] = b0;
                m.qp[f][QP - 1] = p0;
            }
            const float r0 = m.qr[0];
#pragma unroll
            for (int k = 0; k < QP - 1; ++k) m.qr[k] = m.qr[k + 1];
            m.qr[QP - 1] = r0;
        }
    }
}

// One launch of step kernel 13: a tile is NW waves side by side, each wave
// marching its own 56 columns.  A wave whose columns pass q_interior_tile
// takes the interior path (no bounds or CPML code; with K1, every shot of the
// batch has dx == dz and the stencil weights are literals), the others the
// edge path.  Both paths allocate 4 waves per SIMD (at most 128 VGPRs), so
// one launch costs the interior no occupancy, and the edge waves, which are
// few and long, run beside interior ones instead of in a launch of their own.
template <int NW, bool K1>
__global__ __launch_bounds__(BX * NW) void k_step2q(const ShotDev* __restrict__ shots,
                                                     const TileRef* __restrict__ tiles,
                                                     const int* __restrict__ base, int rel, int seg)
{
    const TileRef tr = tiles[blockIdx.x];
    if (tr.shot < 0) return;
    const ShotDev& d = shots[tr.shot];
    const int s = *base + rel - d.t0;
    if (s < 0 || s >= d.nt) return;   // uniform over the block
    const int tj = tr.tile / d.ntx;
    const int ti = tr.tile - tj * d.ntx;
    const int nx = d.nx, nz = d.nz;
    const int zA = tj * seg, zB = min(zA + seg, nz);
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int x0 = (ti * NW + wv) * XQ;
    if (x0 >= nx) return;   // no barriers in this kernel: a wave may leave early
    if (q_interior_tile(nx, nz, d.npml, x0, zA, XQ, seg))
        q_march<2, false, K1>(d, s, x0, zA, zB, threadIdx.x);
    else
        q_march<2, true, false>(d, s, x0, zA, zB, threadIdx.x);
}

// ------------------------------------ passes around a kernel-13 launch
// Level s + k of a shot after a kernel-13 launch of step s, field f (the
// level mapping of q_march).
__device__ __forceinline__ const float* q_level(const ShotDev& d, int s, int k, int f)
{
    const int l = d.mode == MODE_REC ? (d.ntf - 1 - s - k) & 3 : (s + k) & 3;
    return d.f[l][f];
}

__device__ __forceinline__ long q_at(const ShotDev& d, int j, int i)
{
    return (long)(j + MNAD_ZO) * d.pitch + (i + MNAD_XO);
}

// Points per thread per iteration of the passes' row loops: a thread issues
// every load of its QPT points before its first store, so that a wave keeps
// several rows' worth of loads in flight.
constexpr int QPT = 4;

// Plane work of step q on the interior rows blockIdx.x, blockIdx.x +
// gridDim.x, ... of a shot (the workgroups of blockIdx.y = shot).  KIND:
// 1 forward d2u (illumination, and the snapshot when the shot keeps them),
// 2 adjoint imaging, 3 the reconstruction's hand-off slot (snap_step's
// kinds); un, uc, up: levels q + 1, q, q - 1; (sj, si, amp): the source point
// and the source term the step added there.  Per point the arithmetic is
// snap_image_w's.
template <int KIND>
__device__ __forceinline__ void q_plane(const ShotDev& d, const SnapStep& st, const float* un,
                                        const float* uc, const float* up, int sj, int si,
                                        float amp)
{
    const int np = d.npml, nxi = d.nx - 2 * np, nzi = d.nz - 2 * np;
    if (nxi <= 0 || nzi <= 0) return;
    const bool f32 = d.sfmt == SNAP_F32 || d.sfmt == SNAP_SLOT, f16 = d.sfmt == SNAP_F16;
    const int t = threadIdx.x;
    for (int jj = blockIdx.x; jj < nzi; jj += gridDim.x) {
        const long o0 = q_at(d, np + jj, np);
        const long qrow = (long)jj * nxi;
        const int isrc = np + jj == sj ? si - np : -1;   // the source column on this row
        const gfloat* a = gp(un) + o0;
        const gfloat* b = gp(uc) + o0;
        const gfloat* c = gp(up) + o0;
        for (int i0 = t; i0 < nxi; i0 += 256 * QPT) {
            float vn[QPT], vc[QPT], vp[QPT], x0[QPT], x1[QPT];
#pragma unroll
            for (int k = 0; k < QPT; ++k) {
                const int ii = i0 + 256 * k;
                if (ii >= nxi) continue;
                vn[k] = a[ii];
                if (KIND == 2) {
                    if (f32) {
                        x0[k] = (gp(static_cast<const float*>(st.plane)) + qrow)[ii];
                    } else {
                        const unsigned short h = ((const gu16*)st.plane + qrow)[ii];
                        x0[k] = f16 ? __half2float(__ushort_as_half(h)) * d.sinv
                                    : __bfloat162float(__ushort_as_bfloat16(h));
                    }
                    vc[k] = (gp(d.img) + qrow)[ii];
                    vp[k] = (gp(d.imgc) + qrow)[ii];
                } else {
                    vc[k] = b[ii];
                    vp[k] = c[ii];
                    if (KIND == 1) x1[k] = (gp(d.illum) + qrow)[ii];
                }
            }
#pragma unroll
            for (int k = 0; k < QPT; ++k) {
                const int ii = i0 + 256 * k;
                if (ii >= nxi) continue;
                if (KIND == 2) {
                    const float x = x0[k] * vn[k];
                    const float sum = vc[k], cc = vp[k];
                    const float y = x - cc;
                    const float tt = sum + y;
                    (gp(d.imgc) + qrow)[ii] = (tt - sum) - y;
                    (gp(d.img) + qrow)[ii] = tt;
                    continue;
                }
                float du = (vn[k] - 2.0f * vc[k]) + vp[k];
                if (ii == isrc) du -= amp;
                const float d2u = du * d.inv_dt2;
                if (KIND == 3) {
                    (gp(static_cast<float*>(st.plane)) + qrow)[ii] = d2u;
                    continue;
                }
                (gp(d.illum) + qrow)[ii] = x1[k] + d2u * d2u;
                if (st.plane) {
                    if (f32) {
                        (gp(static_cast<float*>(st.plane)) + qrow)[ii] = d2u;
                    } else {
                        const unsigned short h16 = __half_as_ushort(__float2half(d2u * d.sscale));
                        const unsigned short b16 = __bfloat16_as_ushort(__float2bfloat16(d2u));
                        ((gu16*)st.plane + qrow)[ii] = f16 ? h16 : b16;
                    }
                }
            }
        }
    }
}

// After a launch of steps s, s+1 of every shot, off the stored levels; the
// workgroups of blockIdx.y = k work on shot k, so that every shot's work runs
// at once.  PASS 0: snapshots and illumination (forward), imaging (adjoint),
// gathers and ring codes (forward).  PASS 1 (launched after PASS 0): the
// reconstruction's hand-off slot, so that an adjoint reads a slot before the
// reconstruction overwrites it.  The increment of step q at a point is the
// stored second difference of the levels, (U(q+1) - 2 U(q)) + U(q-1), less
// the source term at the source point: the update's own increment up to
// float rounding of the levels (the reverse pass and the snapshots need it to
// about 1e-5, the image tolerance being 1e-3).
template <int PASS>
__global__ __launch_bounds__(256) void k_q_post(const ShotDev* __restrict__ shots,
                                                const int* __restrict__ base, int rel)
{
    const ShotDev& d = shots[blockIdx.y];
    const int s = *base + rel - d.t0;
    if (s < 0 || s >= d.nt) return;   // uniform
    const bool second = s + 1 < d.nt;
    for (int h = 0; h < (second ? 2 : 1); ++h) {
        const int q = s + h;   // step q turns level q into level q+1
        const SnapStep st = snap_step(d, q);
        if (PASS == 0 ? !(st.kind == 1 || st.kind == 2) : st.kind != 3) continue;
        const float* un = q_level(d, s, h + 1, 0);
        const float* uc = q_level(d, s, h, 0);
        const float* up = q_level(d, s, h - 1, 0);
        int sj = -1, si = -1;
        float amp = 0.0f;
        if (d.src_j >= 0 && d.mode != MODE_ADJ) {
            sj = d.src_j;
            si = d.src_i;
            amp = d.mode == MODE_REC ? d.src_amp[d.ntf - q] : d.src_amp[q + 1];
        }
        if (PASS == 1) q_plane<3>(d, st, un, uc, up, sj, si, amp);
        else if (st.kind == 1) q_plane<1>(d, st, un, uc, up, sj, si, amp);
        else q_plane<2>(d, st, un, uc, up, sj, si, amp);
    }
    if (PASS != 0 || d.mode != MODE_FWD) return;
    const int np = d.npml;
    const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nth = (long)gridDim.x * blockDim.x;
    if (d.nrec > 0) {   // gathers: step q records level q at each receiver
        const float* l0 = q_level(d, s, 0, 0);
        const float* l1 = q_level(d, s, 1, 0);
        for (int j = d.rlo; j <= d.rhi; ++j) {
            const int off = d.rrow[j];
            if (off < 0) continue;
            for (long i = tid; i < d.nx; i += nth) {
                const int r = d.rcol[off + i];
                if (r < 0) continue;
                const long o = q_at(d, j, (int)i);
                d.gather[(long)s * d.nrec + r] = l0[o];
                if (second) d.gather[(long)(s + 1) * d.nrec + r] = l1[o];
            }
        }
    }
    if (d.rord > 0) {   // ring codes of levels s+1 and s+2, in order per cell
        for (long r = tid; r < d.nring; r += nth) {
            int j, i;
            ring_point((int)r, d.nz, d.nx, np, &j, &i);
            const long o = q_at(d, j, i);
            ring_put(d, s + 1, (unsigned)r, q_level(d, s, 1, 0)[o], q_level(d, s, 1, 1)[o],
                     q_level(d, s, 1, 2)[o]);
            if (second)
                ring_put(d, s + 2, (unsigned)r, q_level(d, s, 2, 0)[o], q_level(d, s, 2, 1)[o],
                         q_level(d, s, 2, 2)[o]);
        }
    }
}

// Before a launch of steps s, s+1: a reconstruction's ring cells get levels
// c-1 and c-2 (c = ntf - 1 - s) decoded from the saved ring, in order per
// cell (rec_one's decode, two levels at a time); blockIdx.y = shot.
__global__ __launch_bounds__(256) void k_q_pre(const ShotDev* __restrict__ shots,
                                               const int* __restrict__ base, int rel)
{
    const ShotDev& d = shots[blockIdx.y];
    if (d.mode != MODE_REC || d.rord <= 0) return;
    const int s = *base + rel - d.t0;
    if (s < 0 || s >= d.nt) return;
    const bool second = s + 1 < d.nt;
    const int c = d.ntf - 1 - s;
    const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nth = (long)gridDim.x * blockDim.x;
    for (long r = tid; r < d.nring; r += nth) {
        int j, i;
        ring_point((int)r, d.nz, d.nx, d.npml, &j, &i);
        const long o = q_at(d, j, i);
        float v[3];
        ring_get(d, c + 1, (unsigned)r, v);
#pragma unroll
        for (int f = 0; f < 3; ++f) ((float*)q_level(d, s, 1, f))[o] = v[f];
        if (second) {
            ring_get(d, c, (unsigned)r, v);
#pragma unroll
            for (int f = 0; f < 3; ++f) ((float*)q_level(d, s, 2, f))[o] = v[f];
        }
    }
}

__global__ void k_advance(int* base, int n)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *base += n;
}

// Initial state: U at the source holds the first injection a_src w(1).
__global__ void k_init_source(ShotDev* shots, int nshots)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nshots) return;
    const ShotDev& d = shots[k];
    if (d.mode != MODE_FWD || d.src_j < 0) return;
    d.f[0][0][(long)(d.src_j + MNAD_ZO) * d.pitch + (d.src_i + MNAD_XO)] = d.src_amp[0];
}

// radj[s][u] = a_u * (dobs[nt-s][u] - dsyn[nt-s][u]) for s >= 1, 0 for s = 0.
__global__ void k_make_radj(const float* __restrict__ dobs, const float* __restrict__ dsyn,
                            const float* __restrict__ arec, int nt, int nrec,
                            float* __restrict__ radj)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)nt * nrec) return;
    const int s = (int)(k / nrec);
    const int u = (int)(k - (long)s * nrec);
    float val = 0.0f;
    if (s >= 1) {
        const long src = (long)(nt - s) * nrec + u;
        val = arec[u] * (dobs[src] - (dsyn ? dsyn[src] : 0.0f));
    }
    radj[k] = val;
}

// Compensated accumulation of one shot's image and illumination into the
// stack: img_sum += T*(img - imgc), ill_sum += illum (illum already holds the
// un-scaled sum; the T factor is applied here, fwd_PDE_exact_MNAD.m:162).
__global__ void k_stack(const float* __restrict__ img, const float* __restrict__ imgc,
                        const float* __restrict__ illum, float T, long n,
                        float* __restrict__ isum, float* __restrict__ isumc,
                        float* __restrict__ lsum, float* __restrict__ lsumc)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    {
        const float x = T * (img[q] - imgc[q]);
        const float s = isum[q], c = isumc[q];
        const float y = x - c, t = s + y;
        isumc[q] = (t - s) - y;
        isum[q] = t;
    }
    if (illum) {
        const float x = T * illum[q];
        const float s = lsum[q], c = lsumc[q];
        const float y = x - c, t = s + y;
        lsumc[q] = (t - s) - y;
        lsum[q] = t;
    }
}

__global__ void k_fill(float* __restrict__ p, long n, float val)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = val;
}

// ------------------------------------------------------------ host launchers
extern "C" {

hipError_t mnad_launch_setup_model(const double* v, int nz, int nx, int pitch, int npml,
                                   double dx, double dz, double dt, float* rx, float* cfx,
                                   float* cfz, hipStream_t st)
{
    dim3 blk(64, 4);
    dim3 grd((nx + 63) / 64, (nz + 3) / 4);
    hipLaunchKernelGGL(k_setup_model, grd, blk, 0, st, v, nz, nx, pitch, npml, dx, dz, dt,
                       1e-4, rx, cfx, cfz);
    return hipGetLastError();
}

hipError_t mnad_launch_step(const ShotDev* shots, const TileRef* tiles, int ntiles,
                            const int* base, int rel, int kernel, int rows, int brows,
                            hipStream_t st)
{
    const dim3 g(ntiles), b(BX, BZ);
    if (kernel == 0 && brows == 4 && rows == 1)
        hipLaunchKernelGGL((k_step<1, 4>), g, b, 0, st, shots, tiles, base, rel);
    else if (kernel == 0 && brows == 4 && rows == 2)
        hipLaunchKernelGGL((k_step<2, 4>), g, b, 0, st, shots, tiles, base, rel);
    else if (kernel == 0 && brows == 4 && rows == 4)
        hipLaunchKernelGGL((k_step<4, 4>), g, b, 0, st, shots, tiles, base, rel);
    else if (kernel == 0 && brows == 1 && rows == 1)
        hipLaunchKernelGGL((k_step<1, 1>), g, dim3(BX, 1), 0, st, shots, tiles, base, rel);
    else if (kernel == 0 && brows == 2 && rows == 1)
        hipLaunchKernelGGL((k_step<1, 2>), g, dim3(BX, 2), 0, st, shots, tiles, base, rel);
    else if (kernel == 0 && brows == 8 && rows == 1)
        hipLaunchKernelGGL((k_step<1, 8>), g, dim3(BX, 8), 0, st, shots, tiles, base, rel);
    else if (kernel == 0 && brows == 16 && rows == 1)
        hipLaunchKernelGGL((k_step<1, 16>), g, dim3(BX, 16), 0, st, shots, tiles, base, rel);
    else if (kernel == 4 && brows == 1 && rows == 1)
        hipLaunchKernelGGL((k_stepw<1>), g, dim3(BX, 1), 0, st, shots, tiles, base, rel);
    else if (kernel == 4 && brows == 2 && rows == 1)
        hipLaunchKernelGGL((k_stepw<2>), g, dim3(BX, 2), 0, st, shots, tiles, base, rel);
    else if (kernel == 4 && brows == 4 && rows == 1)
        hipLaunchKernelGGL((k_stepw<4>), g, b, 0, st, shots, tiles, base, rel);
    else if (kernel == 4 && brows == 8 && rows == 1)
        hipLaunchKernelGGL((k_stepw<8>), g, dim3(BX, 8), 0, st, shots, tiles, base, rel);
    else if (kernel == 6 && brows == 1 && rows == 1)
        hipLaunchKernelGGL(k_stepx<1>, g, dim3(BX, 1), 0, st, shots, tiles, base, rel);
    else if (kernel == 6 && brows == 2 && rows == 1)
        hipLaunchKernelGGL(k_stepx<2>, g, dim3(BX, 2), 0, st, shots, tiles, base, rel);
    else if (kernel == 6 && brows == 4 && rows == 1)
        hipLaunchKernelGGL(k_stepx<4>, g, b, 0, st, shots, tiles, base, rel);
    else if (kernel == 11 && brows == 1 && rows == 2)
        hipLaunchKernelGGL(k_stepw2, g, dim3(BX, 1), 0, st, shots, tiles, base, rel);
    else if (kernel == 10 && brows == 1 && rows == 1)
        hipLaunchKernelGGL(k_stepx2<1>, g, dim3(BX, 1), 0, st, shots, tiles, base, rel);
    else if (kernel == 10 && brows == 2 && rows == 1)
        hipLaunchKernelGGL(k_stepx2<2>, g, dim3(BX, 2), 0, st, shots, tiles, base, rel);
    else if (kernel == 10 && brows == 4 && rows == 1)
        hipLaunchKernelGGL(k_stepx2<4>, g, b, 0, st, shots, tiles, base, rel);
    else if (kernel == 7 && brows == 4 && rows == 3)
        hipLaunchKernelGGL((k_step2w<12, 4>), g, b, 0, st, shots, tiles, base, rel);
    else if (kernel == 7 && brows == 8 && rows == 3)
        hipLaunchKernelGGL((k_step2w<12, 8>), g, dim3(BX, 8), 0, st, shots, tiles, base, rel);
    else if (kernel == 7 && brows == 16 && rows == 3)
        hipLaunchKernelGGL((k_step2w<12, 16>), g, dim3(BX, 16), 0, st, shots, tiles, base, rel);
    else if (kernel == 7 && brows == 4 && rows == 5)
        hipLaunchKernelGGL((k_step2w<20, 4>), g, b, 0, st, shots, tiles, base, rel);
    else if (kernel == 7 && brows == 8 && rows == 5)
        hipLaunchKernelGGL((k_step2w<20, 8>), g, dim3(BX, 8), 0, st, shots, tiles, base, rel);
    else if (kernel == 7 && brows == 8 && rows == 7)
        hipLaunchKernelGGL((k_step2w<28, 8>), g, dim3(BX, 8), 0, st, shots, tiles, base, rel);
    else if (kernel == 7 && brows == 16 && rows == 7)
        hipLaunchKernelGGL((k_step2w<28, 16>), g, dim3(BX, 16), 0, st, shots, tiles, base, rel);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// Resident kernel: block heights 4 .. 64.
static const void* res_kernel(int bh)
{
    if (bh < 4 || bh > 64) return nullptr;
    return reinterpret_cast<const void*>(&k_res);
}

static size_t res_lds(int bh)
{
    return ((size_t)2 * 3 * (bh + 4) * RES_LW + (size_t)bh * BX) * sizeof(float);
}

int mnad_resident_max_rows() { return 64; }

// Blocks of height bh that one CU holds at once.
hipError_t mnad_resident_fit(int bh, int* per_cu)
{
    const void* k = res_kernel(bh);
    if (!k) return hipErrorInvalidValue;
    // Allow the largest block (64 rows, above the default 64 KiB dynamic-LDS
    // limit of a launch) once and for all: the same value on every call, so
    // contexts on other threads cannot lower it between this check and their
    // launch.
    const hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)res_lds(64));
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k, BX * RES_NW, res_lds(bh));
}

// 64-bit edge records one block of height bh exchanges (both slots, three fields).
long mnad_resident_edge_records(int bh) { return 2L * 3 * res_edge_cells(bh); }

// Every step of a batch in one cooperative launch (all blocks resident);
// ebuf (nblocks * mnad_resident_edge_records(bh) records) and *err must be zero.
hipError_t mnad_launch_resident(const ShotDev* shots, const ResBlock* blocks, int nblocks,
                                unsigned long long* ebuf, int* err, int nsteps, int bh, int spin,
                                hipStream_t st)
{
    const void* k = res_kernel(bh);
    if (!k || nblocks <= 0 || spin < 0) return hipErrorInvalidValue;
    void* args[] = {&shots, &blocks, &ebuf, &err, &nsteps, &bh, &spin};
    return hipLaunchCooperativeKernel(k, dim3(nblocks), dim3(BX, RES_NW), args, res_lds(bh), st);
}

// Step kernel 13 over the tile table, two steps (s, s+1) per launch.
// passes: the batch's work around the step (bit 0: k_q_pre before it; bits 1
// and 2: k_q_post<0> and k_q_post<1> after it).
hipError_t mnad_launch_step2q(const ShotDev* shots, int nshots, const TileRef* tiles, int ntiles,
                              const int* base, int rel, int rows, int brows, int k1, int passes,
                              hipStream_t st)
{
    if (rows < 1 || ntiles < 0) return hipErrorInvalidValue;
    if (brows != 1 && brows != 2 && brows != 4) return hipErrorInvalidValue;
    const int seg = 8 * rows;
    // passes: blockIdx.y = shot; the workgroups per shot shrink as the batch
    // grows (about 8192 in all, 64 .. 2048 per shot)
    const dim3 pg(std::min(2048, std::max(64, (8192 + nshots - 1) / std::max(nshots, 1))),
                  std::max(nshots, 0)),
        pb(256);
    if (nshots <= 0) passes = 0;
    if (passes & 1) hipLaunchKernelGGL(k_q_pre, pg, pb, 0, st, shots, base, rel);
    if (ntiles > 0) {
#define MNAD_Q2(NW)                                                                            \
    do {                                                                                       \
        if (k1)                                                                                \
            hipLaunchKernelGGL((k_step2q<NW, true>), dim3(ntiles), dim3(BX, NW), 0, st, shots, \
                               tiles, base, rel, seg);                                         \
        else                                                                                   \
            hipLaunchKernelGGL((k_step2q<NW, false>), dim3(ntiles), dim3(BX, NW), 0, st,       \
                               shots, tiles, base, rel, seg);                                  \
    } while (0)
        if (brows == 1) MNAD_Q2(1);
        else if (brows == 2) MNAD_Q2(2);
        else MNAD_Q2(4);
#undef MNAD_Q2
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (passes & 2) hipLaunchKernelGGL(k_q_post<0>, pg, pb, 0, st, shots, base, rel);
    if (passes & 4) hipLaunchKernelGGL(k_q_post<1>, pg, pb, 0, st, shots, base, rel);
    return hipGetLastError();
}

hipError_t mnad_launch_advance(int* base, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, base, n);
    return hipGetLastError();
}

hipError_t mnad_launch_init_source(ShotDev* shots, int nshots, hipStream_t st)
{
    hipLaunchKernelGGL(k_init_source, dim3((nshots + 63) / 64), dim3(64), 0, st, shots, nshots);
    return hipGetLastError();
}

hipError_t mnad_launch_make_radj(const float* dobs, const float* dsyn, const float* arec,
                                 int nt, int nrec, float* radj, hipStream_t st)
{
    const long n = (long)nt * nrec;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_make_radj, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       dobs, dsyn, arec, nt, nrec, radj);
    return hipGetLastError();
}

hipError_t mnad_launch_stack(const float* img, const float* imgc, const float* illum, float T,
                             long n, float* isum, float* isumc, float* lsum, float* lsumc,
                             hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_stack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, img, imgc,
                       illum, T, n, isum, isumc, lsum, lsumc);
    return hipGetLastError();
}

hipError_t mnad_launch_fill(float* p, long n, float val, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, val);
    return hipGetLastError();
}

}  // extern "C"

}  // namespace mnad
