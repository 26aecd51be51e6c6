// One grid row per wave against waves that sweep down R rows, at the
// 561 x 1061 grid, with 0 or 80 extra FMAs per point standing in for the
// increment arithmetic.
// A: kernel-4 pattern (3x3 +-2 neighbourhood of U, V, W per point, one row of 64 per wave).
// B: sweep: a wave owns 60 columns x R rows, keeps a register window of rows,
//    takes column neighbours by DPP, prefetches PD rows ahead.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdio>
#include <vector>
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("err %s line %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)
typedef __attribute__((address_space(1))) float gfloat;

struct Args { const float *U, *V, *W, *R; float *PU, *PV, *PW; const int* table; int nz, nx; long P; int ntx; };

__device__ __forceinline__ float lp(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float ln(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x130, 0xf, 0xf, false));
}

template <int NV>
__device__ __forceinline__ float work(float a, float b, float c)
{
    float x = a, y = b;
#pragma unroll
    for (int k = 0; k < NV / 2; ++k) { x = fmaf(x, 0.999f, y * 1e-3f); y = fmaf(y, 0.998f, c); }
    return x + y;
}

template <int NV>
__global__ __launch_bounds__(64) void k_row(Args a)
{
    const int t = a.table[blockIdx.x];
    if (t < 0) return;
    const int j = t / a.ntx, ti = t - j * a.ntx;
    const int i0 = ti * 64;
    const unsigned lane = threadIdx.x;
    if (i0 + (int)lane >= a.nx) return;
    const long P = a.P;
    const long org = (long)(j + 2) * P + i0 + 32;
    const gfloat* u = (const gfloat*)a.U + org + lane; const gfloat* v = (const gfloat*)a.V + org + lane; const gfloat* w = (const gfloat*)a.W + org + lane;
    float acc = 0, acc2 = 0;
#pragma unroll
    for (int r = -1; r <= 1; ++r) {
        acc += u[r * P - 1] * 0.1f + u[r * P] * 0.2f + u[r * P + 1] * 0.3f;
        acc2 += v[r * P - 1] * 0.4f + v[r * P] * 0.5f + v[r * P + 1] * 0.6f;
        acc += w[r * P - 1] * 0.7f + w[r * P] * 0.8f + w[r * P + 1] * 0.9f;
    }
    acc2 += v[-2] * 0.11f + v[2] * 0.12f + w[-2 * P] * 0.13f + w[2 * P] * 0.14f;
    const float rr = ((const gfloat*)a.R + org)[lane];
    float res = work<NV>(acc, acc2, rr);
    gfloat* pu = (gfloat*)a.PU + org; gfloat* pv = (gfloat*)a.PV + org; gfloat* pw = (gfloat*)a.PW + org;
    const float x = pu[lane], y = pv[lane], z = pw[lane];
    pu[lane] = res - x; pv[lane] = res - y; pw[lane] = res - z;
}

// sweep: tile t = tb * ntx60 + tx; rows [tb R, tb R + R), columns [60 tx, 60 tx + 60)
template <int R, int PD, int NV>
__global__ __launch_bounds__(64) void k_sweep(Args a)
{
    const int t = a.table[blockIdx.x];
    if (t < 0) return;
    const int tb = t / a.ntx, tx = t - tb * a.ntx;
    const int j0 = tb * R;
    const unsigned lane = threadIdx.x;
    const int i = tx * 60 - 2 + (int)lane;
    const int ic = min(i, a.nx + 1);
    const bool out = lane >= 2 && lane < 62 && i < a.nx;
    const long P = a.P;
    const gfloat* U = (const gfloat*)a.U + 32 + ic;
    const gfloat* V = (const gfloat*)a.V + 32 + ic;
    const gfloat* W = (const gfloat*)a.W + 32 + ic;
    const gfloat* RR = (const gfloat*)a.R + 32 + ic;
    gfloat* PU = (gfloat*)a.PU + 32 + ic; gfloat* PV = (gfloat*)a.PV + 32 + ic; gfloat* PW = (gfloat*)a.PW + 32 + ic;
    const int rmax = a.nz + 1;   // last loadable grid row (storage row nz + 3)
    auto row = [&](int r) -> long { return (long)(min(r, rmax) + 2) * P; };
    float uC[R + 2], vC[R + 2], wC[R + 4], rx[R], pu[R], pv[R], pw[R];
    // window: U/V rows j0-1, j0; W rows j0-2 .. j0+1
    uC[0] = U[row(j0 - 1)]; uC[1] = U[row(j0)];
    vC[0] = V[row(j0 - 1)]; vC[1] = V[row(j0)];
    wC[0] = W[row(j0 - 2)]; wC[1] = W[row(j0 - 1)]; wC[2] = W[row(j0)]; wC[3] = W[row(j0 + 1)];
#pragma unroll
    for (int k = 0; k < PD && k < R; ++k) {
        uC[k + 2] = U[row(j0 + k + 1)]; vC[k + 2] = V[row(j0 + k + 1)]; wC[k + 4] = W[row(j0 + k + 2)];
        rx[k] = RR[row(j0 + k)]; pu[k] = PU[row(j0 + k)]; pv[k] = PV[row(j0 + k)]; pw[k] = PW[row(j0 + k)];
    }
    float uL[R + 2], uR[R + 2], vL[R + 2], vR[R + 2], wL[R + 4], wR[R + 4];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int kk = k + PD;
        if (kk < R) {
            uC[kk + 2] = U[row(j0 + kk + 1)]; vC[kk + 2] = V[row(j0 + kk + 1)]; wC[kk + 4] = W[row(j0 + kk + 2)];
            rx[kk] = RR[row(j0 + kk)]; pu[kk] = PU[row(j0 + kk)]; pv[kk] = PV[row(j0 + kk)]; pw[kk] = PW[row(j0 + kk)];
        }
        __builtin_amdgcn_sched_barrier(0);
        if (k == 0) {
            uL[0] = lp(uC[0]); uR[0] = ln(uC[0]); uL[1] = lp(uC[1]); uR[1] = ln(uC[1]);
            vL[0] = lp(vC[0]); vR[0] = ln(vC[0]); vL[1] = lp(vC[1]); vR[1] = ln(vC[1]);
            wL[1] = lp(wC[1]); wR[1] = ln(wC[1]); wL[2] = lp(wC[2]); wR[2] = ln(wC[2]);
        }
        uL[k + 2] = lp(uC[k + 2]); uR[k + 2] = ln(uC[k + 2]);
        vL[k + 2] = lp(vC[k + 2]); vR[k + 2] = ln(vC[k + 2]);
        wL[k + 3] = lp(wC[k + 3]); wR[k + 3] = ln(wC[k + 3]);
        const float vww = lp(vL[k + 1]), vee = ln(vR[k + 1]);
        float acc = 0, acc2 = 0;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            acc += uL[k + r] * 0.1f + uC[k + r] * 0.2f + uR[k + r] * 0.3f;
            acc2 += vL[k + r] * 0.4f + vC[k + r] * 0.5f + vR[k + r] * 0.6f;
            acc += wL[k + 1 + r] * 0.7f + wC[k + 1 + r] * 0.8f + wR[k + 1 + r] * 0.9f;
        }
        acc2 += vww * 0.11f + vee * 0.12f + wC[k] * 0.13f + wC[k + 4] * 0.14f;
        const float res = work<NV>(acc, acc2, rx[k]);
        const int j = j0 + k;
        if (out && j < a.nz) {
            const long o = row(j);
            PU[o] = res - pu[k]; PV[o] = res - pv[k]; PW[o] = res - pw[k];
        }
    }
}

int main()
{
    const int nz = 561, nx = 1061; const long P = 1120; const long rows = nz + 12;
    const long n = rows * P;
    std::vector<float*> f(7);
    for (auto& p : f) { CK(hipMalloc(&p, n * 4)); CK(hipMemset(p, 0, n * 4)); }
    hipStream_t sm; CK(hipStreamCreate(&sm));
    hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
    // table: tiles in row-major order, cut into 8 equal runs (XCD x = b % 8)
    auto make_table = [&](int ntiles) {
        const int per = (ntiles + 7) / 8;
        std::vector<int> t(per * 8, -1);
        for (int q = 0; q < ntiles; ++q) { const int x = q / per, m = q - x * per; t[m * 8 + x] = q; }
        return t;
    };
    auto run = [&](const char* name, void (*kern)(Args), int ntx, int ntiles) -> int {
        std::vector<int> table = make_table(ntiles);
        int* dtab; CK(hipMalloc(&dtab, table.size() * 4)); CK(hipMemcpy(dtab, table.data(), table.size() * 4, hipMemcpyHostToDevice));
        hipGraph_t g; hipGraphExec_t ge;
        CK(hipStreamBeginCapture(sm, hipStreamCaptureModeThreadLocal));
        for (int r = 0; r < 64; ++r) {
            const int cb = r & 1;
            Args a{f[cb ? 3 : 0], f[cb ? 4 : 1], f[cb ? 5 : 2], f[6], f[cb ? 0 : 3], f[cb ? 1 : 4], f[cb ? 2 : 5], dtab, nz, nx, P, ntx};
            hipLaunchKernelGGL(kern, dim3((unsigned)table.size()), dim3(64), 0, sm, a);
        }
        CK(hipStreamEndCapture(sm, &g));
        CK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
        for (int k = 0; k < 5; ++k) CK(hipGraphLaunch(ge, sm));
        CK(hipEventRecord(e0, sm));
        const int reps = 50;
        for (int k = 0; k < reps; ++k) CK(hipGraphLaunch(ge, sm));
        CK(hipEventRecord(e1, sm));
        CK(hipEventSynchronize(e1));
        float ms; CK(hipEventElapsedTime(&ms, e0, e1));
        printf("%-32s blocks=%6zu  %.3f us/launch\n", name, table.size(), ms * 1e3 / (reps * 64));
        fflush(stdout);
        CK(hipGraphExecDestroy(ge)); CK(hipGraphDestroy(g)); CK(hipFree(dtab));
        return 0;
    };
    const int ntx64 = (nx + 63) / 64, ntx60 = (nx + 59) / 60;
    int rc = 0;
    for (int rep = 0; rep < 2; ++rep) {
        rc |= run("row, VALU 0", k_row<0>, ntx64, nz * ntx64);
        rc |= run("row, VALU 80", k_row<80>, ntx64, nz * ntx64);
        rc |= run("sweep R4 PD1, VALU 0", k_sweep<4, 1, 0>, ntx60, (nz + 3) / 4 * ntx60);
        rc |= run("sweep R4 PD2, VALU 0", k_sweep<4, 2, 0>, ntx60, (nz + 3) / 4 * ntx60);
        rc |= run("sweep R8 PD2, VALU 0", k_sweep<8, 2, 0>, ntx60, (nz + 7) / 8 * ntx60);
        rc |= run("sweep R8 PD3, VALU 0", k_sweep<8, 3, 0>, ntx60, (nz + 7) / 8 * ntx60);
        rc |= run("sweep R4 PD2, VALU 80", k_sweep<4, 2, 80>, ntx60, (nz + 3) / 4 * ntx60);
        rc |= run("sweep R8 PD2, VALU 80", k_sweep<8, 2, 80>, ntx60, (nz + 7) / 8 * ntx60);
        rc |= run("sweep R8 PD3, VALU 80", k_sweep<8, 3, 80>, ntx60, (nz + 7) / 8 * ntx60);
        rc |= run("sweep R16 PD3, VALU 80", k_sweep<16, 3, 80>, ntx60, (nz + 15) / 16 * ntx60);
    }
    return rc;
}
