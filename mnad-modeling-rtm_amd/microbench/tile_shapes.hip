// Tile shapes for the same neighbourhood loads, XCD-run order, one-shot
// launches at the 561 x 1061 grid: block shapes of one-point waves, waves
// that march down 2 or 4 rows, and lanes holding 2 or 4 columns (16-byte
// loads, neighbours by DPP lane shifts).
#include <hip/hip_runtime.h>
#include <cstdio>
#include <vector>
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("err %s line %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)
typedef __attribute__((address_space(1))) float gfloat;
typedef float vf4 __attribute__((ext_vector_type(4)));
typedef float vf2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) vf4 gfloat4;
typedef __attribute__((address_space(1))) vf2 gfloat2;
struct Args { const float *U, *V, *W, *R; float *PU, *PV, *PW; int nz, nx; long P; int ntx, ntiles, per; };
#define TILE_PROLOGUE \
    const int t_ = (blockIdx.x & 7) * a.per + (blockIdx.x >> 3); \
    if (t_ >= a.ntiles) return; \
    const int tj = t_ / a.ntx, ti = t_ - tj * a.ntx;
__device__ __forceinline__ float nbh_point(const gfloat* u, const gfloat* v, const gfloat* w, long P)
{
    float acc = 0;
#pragma unroll
    for (int r = -1; r <= 1; ++r) {
        const gfloat* p = u + r * P; acc += p[-1] * 0.1f + p[0] * 0.2f + p[1] * 0.3f;
        const gfloat* q = v + r * P; acc += q[-1] * 0.4f + q[0] * 0.5f + q[1] * 0.6f;
        const gfloat* s = w + r * P; acc += s[-1] * 0.7f + s[0] * 0.8f + s[1] * 0.9f;
    }
    acc += v[-2] * 0.11f + v[2] * 0.12f + w[-2 * P] * 0.13f + w[2 * P] * 0.14f;
    return acc;
}

// block = (64*WX) x BR threads: WX waves across x, BR rows; one point per thread
template <int WX, int BR>
__global__ __launch_bounds__(64 * WX * BR) void s_blk(Args a)
{
    TILE_PROLOGUE
    const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int wy = wid / WX, wx = wid - wy * WX;
    const int j = tj * BR + wy;
    const int i0 = ti * 64 * WX + wx * 64; const unsigned lane = threadIdx.x & 63;
    if (j >= a.nz || i0 + (int)lane >= a.nx) return;
    const long P = a.P;
    const long org = (long)(j + 2) * P + i0 + 32;
    float acc = nbh_point((const gfloat*)a.U + org + lane, (const gfloat*)a.V + org + lane, (const gfloat*)a.W + org + lane, P);
    acc *= ((const gfloat*)a.R + org)[lane];
    gfloat* pu = (gfloat*)a.PU + org; gfloat* pv = (gfloat*)a.PV + org; gfloat* pw = (gfloat*)a.PW + org;
    const float x = pu[lane], y = pv[lane], z = pw[lane];
    pu[lane] = acc - x; pv[lane] = acc - y; pw[lane] = acc - z;
}

// z-march: a wave owns 64 columns x RZ rows, keeps the column window in registers
template <int RZ, int BR>
__global__ __launch_bounds__(64 * BR) void s_march(Args a)
{
    TILE_PROLOGUE
    const int wy = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int j0 = (tj * BR + wy) * RZ;
    const int i0 = ti * 64; const unsigned lane = threadIdx.x;
    if (j0 >= a.nz || i0 + (int)lane >= a.nx) return;
    const long P = a.P;
    const long org = (long)(j0 + 2) * P + i0 + 32;
    const gfloat* U = (const gfloat*)a.U + org + lane;
    const gfloat* V = (const gfloat*)a.V + org + lane;
    const gfloat* W = (const gfloat*)a.W + org + lane;
    float u[RZ + 2][3], v[RZ + 2][3], w[RZ + 4], ws[RZ + 2][2], ve[RZ][2];
#pragma unroll
    for (int k = 0; k < RZ + 2; ++k) {
        const gfloat* p = U + (k - 1) * P; u[k][0] = p[-1]; u[k][1] = p[0]; u[k][2] = p[1];
        const gfloat* q = V + (k - 1) * P; v[k][0] = q[-1]; v[k][1] = q[0]; v[k][2] = q[1];
        const gfloat* s = W + (k - 1) * P; ws[k][0] = s[-1]; ws[k][1] = s[1];
    }
#pragma unroll
    for (int k = 0; k < RZ + 4; ++k) w[k] = W[(k - 2) * P];
#pragma unroll
    for (int k = 0; k < RZ; ++k) { ve[k][0] = V[k * P - 2]; ve[k][1] = V[k * P + 2]; }
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
        if (j0 + r >= a.nz) break;
        float acc = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            acc += u[r + k][0] * 0.1f + u[r + k][1] * 0.2f + u[r + k][2] * 0.3f;
            acc += v[r + k][0] * 0.4f + v[r + k][1] * 0.5f + v[r + k][2] * 0.6f;
            acc += ws[r + k][0] * 0.7f + w[r + k + 1] * 0.8f + ws[r + k][1] * 0.9f;
        }
        acc += ve[r][0] * 0.11f + ve[r][1] * 0.12f + w[r] * 0.13f + w[r + 4] * 0.14f;
        const long o = org + r * P;
        acc *= ((const gfloat*)a.R + o)[lane];
        gfloat* pu = (gfloat*)a.PU + o; gfloat* pv = (gfloat*)a.PV + o; gfloat* pw = (gfloat*)a.PW + o;
        const float x = pu[lane], y = pv[lane], z = pw[lane];
        pu[lane] = acc - x; pv[lane] = acc - y; pw[lane] = acc - z;
    }
}

__device__ __forceinline__ float shr1(float x) { return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x138, 0xf, 0xf, false)); }
__device__ __forceinline__ float shl1(float x) { return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x130, 0xf, 0xf, false)); }

// ---- x4: lane holds 4 consecutive columns; wave covers 256 columns, outputs lanes 1..62 (248 cols)
template <int BR>
__global__ __launch_bounds__(64 * BR) void s_x4(Args a)
{
    TILE_PROLOGUE
    const int j = tj * BR + __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int i0 = ti * 248 - 4;
    const unsigned lane = threadIdx.x;
    if (j >= a.nz) return;
    const long P = a.P;
    const long org = (long)(j + 2) * P + i0 + 32;   // 16-byte aligned (i0 + 32 = 248 ti + 28)
    float acc[4] = {0, 0, 0, 0};
    auto row = [&](const float* base, long off, float c0, float c1, float c2) {
        const vf4 e = ((const gfloat4*)((const gfloat*)base + org + off))[lane];
        const float l = shr1(e.w), r = shl1(e.x);
        acc[0] += l * c0 + e.x * c1 + e.y * c2;
        acc[1] += e.x * c0 + e.y * c1 + e.z * c2;
        acc[2] += e.y * c0 + e.z * c1 + e.w * c2;
        acc[3] += e.z * c0 + e.w * c1 + r * c2;
        return e;
    };
#pragma unroll
    for (int r = -1; r <= 1; ++r) {
        row(a.U, r * P, 0.1f, 0.2f, 0.3f);
        const vf4 e = row(a.V, r * P, 0.4f, 0.5f, 0.6f);
        if (r == 0) {
            const float l2 = shr1(e.z), l1 = shr1(e.w), r1 = shl1(e.x), r2 = shl1(e.y);
            acc[0] += l2 * 0.11f + e.z * 0.12f; acc[1] += l1 * 0.11f + e.w * 0.12f;
            acc[2] += e.x * 0.11f + r1 * 0.12f; acc[3] += e.y * 0.11f + r2 * 0.12f;
        }
        row(a.W, r * P, 0.7f, 0.8f, 0.9f);
    }
    {
        const vf4 n = ((const gfloat4*)((const gfloat*)a.W + org - 2 * P))[lane];
        const vf4 s = ((const gfloat4*)((const gfloat*)a.W + org + 2 * P))[lane];
        acc[0] += n.x * 0.13f + s.x * 0.14f; acc[1] += n.y * 0.13f + s.y * 0.14f;
        acc[2] += n.z * 0.13f + s.z * 0.14f; acc[3] += n.w * 0.13f + s.w * 0.14f;
    }
    const int ic = i0 + 4 * (int)lane;
    if (lane < 1 || lane >= 63 || ic >= a.nx) return;
    const vf4 rr = ((const gfloat4*)((const gfloat*)a.R + org))[lane];
    gfloat4* pu = (gfloat4*)((gfloat*)a.PU + org); gfloat4* pv = (gfloat4*)((gfloat*)a.PV + org); gfloat4* pw = (gfloat4*)((gfloat*)a.PW + org);
    const vf4 x = pu[lane], y = pv[lane], z = pw[lane];
    vf4 o;
    o.x = acc[0] * rr.x; o.y = acc[1] * rr.y; o.z = acc[2] * rr.z; o.w = acc[3] * rr.w;
    pu[lane] = (vf4)(o.x - x.x, o.y - x.y, o.z - x.z, o.w - x.w);
    pv[lane] = (vf4)(o.x - y.x, o.y - y.y, o.z - y.z, o.w - y.w);
    pw[lane] = (vf4)(o.x - z.x, o.y - z.y, o.z - z.z, o.w - z.w);
}

// ---- x2: lane holds 2 columns; wave covers 128 columns, outputs lanes 1..62 (124 cols)
template <int BR>
__global__ __launch_bounds__(64 * BR) void s_x2(Args a)
{
    TILE_PROLOGUE
    const int j = tj * BR + __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int i0 = ti * 124 - 2;
    const unsigned lane = threadIdx.x;
    if (j >= a.nz) return;
    const long P = a.P;
    const long org = (long)(j + 2) * P + i0 + 32;
    float acc[2] = {0, 0};
    auto row = [&](const float* base, long off, float c0, float c1, float c2) {
        const vf2 e = ((const gfloat2*)((const gfloat*)base + org + off))[lane];
        const float l = shr1(e.y), r = shl1(e.x);
        acc[0] += l * c0 + e.x * c1 + e.y * c2;
        acc[1] += e.x * c0 + e.y * c1 + r * c2;
        return e;
    };
#pragma unroll
    for (int r = -1; r <= 1; ++r) {
        row(a.U, r * P, 0.1f, 0.2f, 0.3f);
        const vf2 e = row(a.V, r * P, 0.4f, 0.5f, 0.6f);
        if (r == 0) {
            const float l2 = shr1(e.x), l1 = shr1(e.y), r1 = shl1(e.x), r2 = shl1(e.y);
            acc[0] += l2 * 0.11f + r1 * 0.12f; acc[1] += l1 * 0.11f + r2 * 0.12f;
        }
        row(a.W, r * P, 0.7f, 0.8f, 0.9f);
    }
    {
        const vf2 n = ((const gfloat2*)((const gfloat*)a.W + org - 2 * P))[lane];
        const vf2 s = ((const gfloat2*)((const gfloat*)a.W + org + 2 * P))[lane];
        acc[0] += n.x * 0.13f + s.x * 0.14f; acc[1] += n.y * 0.13f + s.y * 0.14f;
    }
    const int ic = i0 + 2 * (int)lane;
    if (lane < 1 || lane >= 63 || ic >= a.nx) return;
    const vf2 rr = ((const gfloat2*)((const gfloat*)a.R + org))[lane];
    gfloat2* pu = (gfloat2*)((gfloat*)a.PU + org); gfloat2* pv = (gfloat2*)((gfloat*)a.PV + org); gfloat2* pw = (gfloat2*)((gfloat*)a.PW + org);
    const vf2 x = pu[lane], y = pv[lane], z = pw[lane];
    const float o0 = acc[0] * rr.x, o1 = acc[1] * rr.y;
    pu[lane] = (vf2)(o0 - x.x, o1 - x.y);
    pv[lane] = (vf2)(o0 - y.x, o1 - y.y);
    pw[lane] = (vf2)(o0 - z.x, o1 - z.y);
}


int main()
{
    const int nz = 561, nx = 1061; const long P = 1120; const long rows = nz + 24;
    const long n = rows * P;
    std::vector<float*> f(7);
    for (auto& p : f) { CK(hipMalloc(&p, n * 4)); CK(hipMemset(p, 0, n * 4)); }
    hipStream_t st; CK(hipStreamCreate(&st));
    hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
    auto run = [&](const char* name, void (*kern)(Args), dim3 blk, int tw, int th) -> int {
        const int ntx = (nx + tw - 1) / tw, ntz = (nz + th - 1) / th, nt = ntx * ntz;
        const int per = (nt + 7) / 8, nb = per * 8;
        hipGraph_t g; hipGraphExec_t ge;
        CK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (int r = 0; r < 64; ++r) {
            const int cb = r & 1;
            Args a{f[cb ? 3 : 0], f[cb ? 4 : 1], f[cb ? 5 : 2], f[6], f[cb ? 0 : 3], f[cb ? 1 : 4], f[cb ? 2 : 5], nz, nx, P, ntx, nt, per};
            hipLaunchKernelGGL(kern, dim3(nb), blk, 0, st, a);
        }
        CK(hipStreamEndCapture(st, &g));
        CK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
        for (int k = 0; k < 5; ++k) CK(hipGraphLaunch(ge, st));
        CK(hipEventRecord(e0, st));
        const int reps = 50;
        for (int k = 0; k < reps; ++k) CK(hipGraphLaunch(ge, st));
        CK(hipEventRecord(e1, st));
        CK(hipEventSynchronize(e1));
        float ms; CK(hipEventElapsedTime(&ms, e0, e1));
        const double us = ms * 1e3 / (reps * 64);
        printf("%-14s tile %4dx%-3d blocks=%5d  %.3f us/launch  %.2f TB/s (23.8 MB)\n", name, tw, th, nb, us, 23.8e6 / us / 1e6);
        CK(hipGraphExecDestroy(ge)); CK(hipGraphDestroy(g));
        return 0;
    };
    int rc = 0;
    rc |= run("blk 1x1", s_blk<1, 1>, dim3(64), 64, 1);
    rc |= run("blk 1x4", s_blk<1, 4>, dim3(256), 64, 4);
    rc |= run("blk 2x2", s_blk<2, 2>, dim3(256), 128, 2);
    rc |= run("blk 1x8", s_blk<1, 8>, dim3(512), 64, 8);
    rc |= run("march 2 br1", s_march<2, 1>, dim3(64, 1), 64, 2);
    rc |= run("march 2 br2", s_march<2, 2>, dim3(64, 2), 64, 4);
    rc |= run("march 2 br4", s_march<2, 4>, dim3(64, 4), 64, 8);
    rc |= run("march 4 br1", s_march<4, 1>, dim3(64, 1), 64, 4);
    rc |= run("march 4 br2", s_march<4, 2>, dim3(64, 2), 64, 8);
    rc |= run("march 4 br4", s_march<4, 4>, dim3(64, 4), 64, 16);
    rc |= run("x2 br1", s_x2<1>, dim3(64, 1), 124, 1);
    rc |= run("x2 br2", s_x2<2>, dim3(64, 2), 124, 2);
    rc |= run("x2 br4", s_x2<4>, dim3(64, 4), 124, 4);
    rc |= run("x4 br1", s_x4<1>, dim3(64, 1), 248, 1);
    rc |= run("x4 br2", s_x4<2>, dim3(64, 2), 248, 2);
    rc |= run("x4 br4", s_x4<4>, dim3(64, 4), 248, 4);
    return rc;
}
