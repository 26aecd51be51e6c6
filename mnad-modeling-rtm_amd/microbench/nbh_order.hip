// Step-kernel access patterns at the 561 x 1061 grid, one-shot launches in a
// hipGraph: the 3x3 (+-2) neighbourhoods of U, Vs, Ws (kernel-4 loads with a
// trivial combination), rows only, columns only and the own point only, each
// with the natural block order and with XCD runs (workgroup b on XCD b % 8
// takes a contiguous run of tiles).
#include <hip/hip_runtime.h>
#include <cstdio>
#include <vector>
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("err %s line %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)
typedef __attribute__((address_space(1))) float gfloat;

struct Args { const float *U, *V, *W, *R; float *PU, *PV, *PW; int nz, nx; long P; int ntx, ntiles, per; };

// tile of block b: natural (b) or XCD runs (XCD x = b % 8 gets tiles [x per, (x+1) per))
template <bool XCD>
__device__ __forceinline__ int tile_of(const Args& a)
{
    if (!XCD) return blockIdx.x;
    const int x = blockIdx.x & 7, m = blockIdx.x >> 3;
    return x * a.per + m;
}

// MODE 0: full nbh; 1: rows only (no column shifts); 2: columns only (own row, +-1, +-2); 3: point only
template <int MODE, bool XCD>
__global__ __launch_bounds__(256) void s_var(Args a)
{
    const int t = tile_of<XCD>(a);
    if (t >= a.ntiles) return;
    const int tj = t / a.ntx, ti = t - tj * a.ntx;
    const int j = tj * 4 + __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int i0 = ti * 64; const unsigned lane = threadIdx.x;
    if (j >= a.nz || i0 + (int)lane >= a.nx) return;
    const long P = a.P;
    const long org = (long)(j + 2) * P + i0 + 32;
    const gfloat* u = (const gfloat*)a.U + org + lane; const gfloat* v = (const gfloat*)a.V + org + lane; const gfloat* w = (const gfloat*)a.W + org + lane;
    float acc = 0;
    if (MODE == 0) {
#pragma unroll
        for (int r = -1; r <= 1; ++r) {
            acc += u[r * P - 1] * 0.1f + u[r * P] * 0.2f + u[r * P + 1] * 0.3f;
            acc += v[r * P - 1] * 0.4f + v[r * P] * 0.5f + v[r * P + 1] * 0.6f;
            acc += w[r * P - 1] * 0.7f + w[r * P] * 0.8f + w[r * P + 1] * 0.9f;
        }
        acc += v[-2] * 0.11f + v[2] * 0.12f + w[-2 * P] * 0.13f + w[2 * P] * 0.14f;
    } else if (MODE == 1) {
#pragma unroll
        for (int r = -1; r <= 1; ++r) acc += u[r * P] * 0.2f + v[r * P] * 0.5f + w[r * P] * 0.8f;
        acc += w[-2 * P] * 0.13f + w[2 * P] * 0.14f;
    } else if (MODE == 2) {
        acc += u[-1] * 0.1f + u[0] * 0.2f + u[1] * 0.3f;
        acc += v[-2] * 0.11f + v[-1] * 0.4f + v[0] * 0.5f + v[1] * 0.6f + v[2] * 0.12f;
        acc += w[-1] * 0.7f + w[0] * 0.8f + w[1] * 0.9f;
    } else {
        acc += u[0] * 0.2f + v[0] * 0.5f + w[0] * 0.8f;
    }
    acc *= ((const gfloat*)a.R + org)[lane];
    gfloat* pu = (gfloat*)a.PU + org; gfloat* pv = (gfloat*)a.PV + org; gfloat* pw = (gfloat*)a.PW + org;
    const float x = pu[lane], y = pv[lane], z = pw[lane];
    pu[lane] = acc - x; pv[lane] = acc - y; pw[lane] = acc - z;
}

int main()
{
    const int nz = 561, nx = 1061; const long P = 1120; const long rows = nz + 12;
    const long n = rows * P;
    std::vector<float*> f(7);
    for (auto& p : f) { CK(hipMalloc(&p, n * 4)); CK(hipMemset(p, 0, n * 4)); }
    hipStream_t st; CK(hipStreamCreate(&st));
    hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
    auto run = [&](const char* name, void (*kern)(Args), bool xcd) -> int {
        const int ntx = (nx + 63) / 64, ntz = (nz + 3) / 4, nt = ntx * ntz;
        const int per = (nt + 7) / 8, nb = xcd ? per * 8 : nt;
        hipGraph_t g; hipGraphExec_t ge;
        CK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (int r = 0; r < 64; ++r) {
            const int cb = r & 1;
            Args a{f[cb ? 3 : 0], f[cb ? 4 : 1], f[cb ? 5 : 2], f[6], f[cb ? 0 : 3], f[cb ? 1 : 4], f[cb ? 2 : 5], nz, nx, P, ntx, nt, per};
            hipLaunchKernelGGL(kern, dim3(nb), dim3(64, 4), 0, st, a);
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
        printf("%-14s xcd=%d blocks=%5d  %.3f us/launch  %.2f TB/s (23.8 MB)\n", name, (int)xcd, nb, us, 23.8e6 / us / 1e6);
        CK(hipGraphExecDestroy(ge)); CK(hipGraphDestroy(g));
        return 0;
    };
    int rc = 0;
    rc |= run("nbh", s_var<0, false>, false); rc |= run("nbh", s_var<0, true>, true);
    rc |= run("rows only", s_var<1, false>, false); rc |= run("rows only", s_var<1, true>, true);
    rc |= run("cols only", s_var<2, false>, false); rc |= run("cols only", s_var<2, true>, true);
    rc |= run("point", s_var<3, false>, false); rc |= run("point", s_var<3, true>, true);
    return rc;
}
