// CPML strip memories on the kernel-4 access pattern (one wave row per
// workgroup, XCD runs of equal estimated work, strip tiles first) at the
// 561 x 1061 grid with 30-point strips: no memories, the nine memories and
// three coefficients as 12 dword planes (the library layout), or packed into
// float4 planes (4 loads and 3 stores instead of 12 and 9).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdio>
#include <vector>
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("err %s line %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)
typedef __attribute__((address_space(1))) float gfloat;
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) f4 gf4;

struct Set { float* p[9]; float* c[3]; f4* p4[2]; float* p1; f4* c4; };
struct Args { const float *U, *V, *W, *R; float *PU, *PV, *PW; const int* table; int nz, nx, np; long P; int ntx; Set X, Z; };

__device__ __forceinline__ float upd(float& p1, float& p2, float& p3, float d, float c1, float c2, float c3)
{
    p3 = c1 * p3 + 0.5f * c2 * d;
    p2 = c1 * p2 + c2 * d + c3 * (d - p3);
    p1 = c1 * p1 + c3 * (2.0f * d - p2);
    return p1;
}

template <int MODE>
__device__ __forceinline__ float strip(const Set& S, long q, float a, float b, float c)
{
    float r = 0.f;
    if (MODE == 1) {
        float p[9], cc[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) cc[k] = ((const gfloat*)S.c[k])[q];
#pragma unroll
        for (int k = 0; k < 9; ++k) p[k] = ((const gfloat*)S.p[k])[q];
        r += upd(p[0], p[1], p[2], a, cc[0], cc[1], cc[2]);
        r += upd(p[3], p[4], p[5], b, cc[0], cc[1], cc[2]);
        r += upd(p[6], p[7], p[8], c, cc[0], cc[1], cc[2]);
#pragma unroll
        for (int k = 0; k < 9; ++k) ((gfloat*)S.p[k])[q] = p[k];
    } else if (MODE == 2) {
        f4 A = ((const gf4*)S.p4[0])[q], B = ((const gf4*)S.p4[1])[q];
        float p8 = ((const gfloat*)S.p1)[q];
        const f4 cc = ((const gf4*)S.c4)[q];
        float p[9] = {A.x, A.y, A.z, A.w, B.x, B.y, B.z, B.w, p8};
        r += upd(p[0], p[1], p[2], a, cc.x, cc.y, cc.z);
        r += upd(p[3], p[4], p[5], b, cc.x, cc.y, cc.z);
        r += upd(p[6], p[7], p[8], c, cc.x, cc.y, cc.z);
        ((gf4*)S.p4[0])[q] = (f4){p[0], p[1], p[2], p[3]};
        ((gf4*)S.p4[1])[q] = (f4){p[4], p[5], p[6], p[7]};
        ((gfloat*)S.p1)[q] = p[8];
    }
    return r;
}

template <int MODE>
__global__ __launch_bounds__(64) void s_strip(Args a)
{
    const int t = a.table[blockIdx.x];
    if (t < 0) return;
    const int j = t / a.ntx, ti = t - j * a.ntx;
    const int i0 = ti * 64;
    const unsigned lane = threadIdx.x;
    const int i = i0 + (int)lane;
    if (i >= a.nx) return;
    const long P = a.P;
    const long org = (long)(j + 2) * P + i0 + 32;
    const gfloat* u = (const gfloat*)a.U + org + lane; const gfloat* v = (const gfloat*)a.V + org + lane; const gfloat* w = (const gfloat*)a.W + org + lane;
    float acc = 0;
#pragma unroll
    for (int r = -1; r <= 1; ++r) {
        acc += u[r * P - 1] * 0.1f + u[r * P] * 0.2f + u[r * P + 1] * 0.3f;
        acc += v[r * P - 1] * 0.4f + v[r * P] * 0.5f + v[r * P + 1] * 0.6f;
        acc += w[r * P - 1] * 0.7f + w[r * P] * 0.8f + w[r * P + 1] * 0.9f;
    }
    acc += v[-2] * 0.11f + v[2] * 0.12f + w[-2 * P] * 0.13f + w[2 * P] * 0.14f;
    const bool inz = j < a.np || j >= a.nz - a.np;
    const bool inx = i < a.np || i >= a.nx - a.np;
    const long q = (long)j * P + i;
    if (MODE != 0) {
        if (inx) acc -= 0.01f * strip<MODE>(a.X, q, u[-1] - u[1], v[-1] - v[1], w[-1] - w[1]);
        if (inz) acc -= 0.01f * strip<MODE>(a.Z, q, u[-P] - u[P], v[-P] - v[P], w[-P] - w[P]);
    }
    acc *= ((const gfloat*)a.R + org)[lane];
    gfloat* pu = (gfloat*)a.PU + org; gfloat* pv = (gfloat*)a.PV + org; gfloat* pw = (gfloat*)a.PW + org;
    const float x = pu[lane], y = pv[lane], z = pw[lane];
    pu[lane] = acc - x; pv[lane] = acc - y; pw[lane] = acc - z;
}

int main()
{
    const int nz = 561, nx = 1061, np = 30; const long P = 1120; const long rows = nz + 12;
    const long n = rows * P, nq = (long)nz * P;
    std::vector<float*> f(7);
    for (auto& p : f) { CK(hipMalloc(&p, n * 4)); CK(hipMemset(p, 0, n * 4)); }
    Set S[2];
    for (auto& s : S) {
        for (auto& p : s.p) { CK(hipMalloc(&p, nq * 4)); CK(hipMemset(p, 0, nq * 4)); }
        for (auto& p : s.c) { CK(hipMalloc(&p, nq * 4)); CK(hipMemset(p, 0, nq * 4)); }
        for (auto& p : s.p4) { CK(hipMalloc(&p, nq * 16)); CK(hipMemset(p, 0, nq * 16)); }
        CK(hipMalloc(&s.p1, nq * 4)); CK(hipMemset(s.p1, 0, nq * 4));
        CK(hipMalloc(&s.c4, nq * 16)); CK(hipMemset(s.c4, 0, nq * 16));
    }
    // tile table: one row of 64 per tile, weighted XCD split, strip tiles first
    const int ntx = (nx + 63) / 64;
    std::vector<int> list; std::vector<double> wt; std::vector<char> st;
    for (int j = 0; j < nz; ++j)
        for (int ti = 0; ti < ntx; ++ti) {
            const int c0 = ti * 64, c1 = std::min(nx, c0 + 64);
            const bool s = j < np || j >= nz - np || c0 < np || c1 > nx - np;
            list.push_back(j * ntx + ti); st.push_back(s); wt.push_back(s ? 2.0 : 1.0);
        }
    double tot = 0; for (double x : wt) tot += x;
    std::vector<std::vector<int>> runs(8); double acc = 0;
    for (size_t q = 0; q < list.size(); ++q) { double mid = acc + 0.5 * wt[q]; acc += wt[q]; runs[std::min(7, (int)(mid * 8 / tot))].push_back((int)q); }
    size_t per = 0; for (auto& r : runs) per = std::max(per, r.size());
    std::vector<int> table(per * 8, -1);
    for (int x = 0; x < 8; ++x) {
        auto& r = runs[x];
        std::stable_partition(r.begin(), r.end(), [&](int q) { return st[q] != 0; });
        for (size_t m = 0; m < r.size(); ++m) table[m * 8 + x] = list[r[m]];
    }
    int* dtab; CK(hipMalloc(&dtab, table.size() * 4)); CK(hipMemcpy(dtab, table.data(), table.size() * 4, hipMemcpyHostToDevice));
    hipStream_t sm; CK(hipStreamCreate(&sm));
    hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
    auto run = [&](const char* name, void (*kern)(Args), int npml) -> int {
        hipGraph_t g; hipGraphExec_t ge;
        CK(hipStreamBeginCapture(sm, hipStreamCaptureModeThreadLocal));
        for (int r = 0; r < 64; ++r) {
            const int cb = r & 1;
            Args a{f[cb ? 3 : 0], f[cb ? 4 : 1], f[cb ? 5 : 2], f[6], f[cb ? 0 : 3], f[cb ? 1 : 4], f[cb ? 2 : 5], dtab, nz, nx, npml, P, ntx, S[0], S[1]};
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
        printf("%-28s %.3f us/launch\n", name, ms * 1e3 / (reps * 64));
        fflush(stdout);
        CK(hipGraphExecDestroy(ge)); CK(hipGraphDestroy(g));
        return 0;
    };
    int rc = 0;
    for (int rep = 0; rep < 2; ++rep) {
        rc |= run("no strip memories", s_strip<0>, np);
        rc |= run("12 dword planes", s_strip<1>, np);
        rc |= run("float4-packed planes", s_strip<2>, np);
    }
    return rc;
}
