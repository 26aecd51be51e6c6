// mnad_noalloc.hip -- mnad_fwd_noalloc: the forward shot of mnad_fwd without
// any device allocation, for callers that cannot allocate (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Everything the step kernels need beyond the caller's inputs and outputs --
// the FP64 velocity copy for the set-up, a/dx^2, the CPML coefficients, both
// time levels of (U, Vs, Ws), the CPML memories, the source amplitudes, the
// receiver tables, the gather, the shot descriptor, the tile table and the
// step counter -- is carved out of one file-scope device array, `scratch`,
// of kScratchBytes.  Each GPU the library runs on holds its own copy until the
// process ends; its device address is taken once per device, with that device
// current, when a context is created on it.  Calls on one GPU share it and
// keep the last call's contents, so every call clears what must start at zero
// (with a fill kernel, not a memset, so the sequence stays capturable); calls
// on one GPU are serialised by a per-device lock inside this process.
// Launches are plain (no graph instantiation, no events).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "mnad.h"
#include "mnad_internal.h"

namespace mnad {

extern "C" {
hipError_t mnad_launch_setup_model(const double* v, int nz, int nx, int pitch, int npml,
                                   double dx, double dz, double dt, float* rx, float* cfx,
                                   float* cfz, hipStream_t st);
hipError_t mnad_launch_step(const ShotDev* shots, const TileRef* tiles, int ntiles,
                            const int* base, int rel, int kernel, int rows, int brows,
                            hipStream_t st);
hipError_t mnad_launch_init_source(ShotDev* shots, int nshots, hipStream_t st);
hipError_t mnad_launch_fill(float* p, long n, float val, hipStream_t st);
}

constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
void* g_scratch[kMaxDevices] = {};   // device address of `scratch` per device
std::mutex g_init_mu;
std::mutex g_dev_mu[kMaxDevices];    // one allocation-free call per device at a time

// Bytes carved from the scratch array, each piece 256-byte aligned.
struct Carve {
    size_t off = 0;
    size_t take(size_t bytes)
    {
        const size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    }
};
}  // namespace

// The address of this device's `scratch` (the host-side name is only a
// stand-in); called by mnad_ctx_create with the device current.
int noalloc_init_device(int device)
{
    if (device < 0 || device >= kMaxDevices) return MNAD_OK;   // the entry point refuses later
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_scratch[device]) return MNAD_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(MNAD_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", device,
                    hipGetErrorString(e));
    g_scratch[device] = p;
    return MNAD_OK;
}

}  // namespace mnad

using namespace mnad;

#define HIPCHK(x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess)                                                          \
            return fail(MNAD_ERR_HIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                           \
    } while (0)

#define CHK(x)              \
    do {                    \
        int rc_ = (x);      \
        if (rc_ != MNAD_OK) \
            return rc_;     \
    } while (0)

// fwd_PDE_exact_MNAD.m:1 with indx = 0, every device buffer in `scratch`.
extern "C" int mnad_fwd_noalloc(mnad_ctx* ctx, const mnad_grid* g, const double* v, int nt,
                                const double* w, int zs, int xs, int nr, const int* zr,
                                const int* xr, double* datau, double* U, double* V, double* W)
{
    if (!ctx) return fail(MNAD_ERR_ARG, "context is NULL");
    std::lock_guard<std::recursive_mutex> lock(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    if (ctx->device >= kMaxDevices || !g_scratch[ctx->device])
        return fail(MNAD_ERR_STATE, "no scratch array on device %d", ctx->device);
    std::lock_guard<std::mutex> dlk(g_dev_mu[ctx->device]);
    char* base = static_cast<char*>(g_scratch[ctx->device]);

    // ---- arguments (as mnad_fwd) ----
    CHK(check_grid(g));
    if (!v) return fail(MNAD_ERR_ARG, "velocity is NULL");
    if (nt < 1) return fail(MNAD_ERR_ARG, "nt must be >= 1");
    if (!w) return fail(MNAD_ERR_ARG, "wavelet is NULL");
    if (nr < 0 || (nr > 0 && (!zr || !xr))) return fail(MNAD_ERR_ARG, "bad receiver arrays");
    const size_t n = (size_t)g->nz * g->nx;
    double vmin, vmax;
    const size_t bad = scan_model(v, n, &vmin, &vmax);
    if (bad < n)
        return fail(MNAD_ERR_ARG, "velocity must be finite and positive (v[%zu] = %g)", bad, v[bad]);
    if (vmax * g->dt / std::min(g->dx, g->dz) > 1.0)
        return fail(MNAD_ERR_ARG, "vmax*dt/min(dx,dz) = %.3f > 1: unstable",
                    vmax * g->dt / std::min(g->dx, g->dz));
    if (zs < 0 || zs >= g->nz || xs < 0 || xs >= g->nx)
        return fail(MNAD_ERR_ARG, "source (%d,%d) outside the %dx%d grid", zs, xs, g->nz, g->nx);

    // one-step kernels only (the two-step and resident kernels need other buffers)
    TunePick k4{kDefaultKernel, kDefaultRows, kDefaultBrows, 1, 0.0, {}};
    const bool other = two_step_kernel(ctx->kernel) || resident_kernel(ctx->kernel);
    TuningGuard tg(ctx, other ? &k4 : nullptr);

    const int np = g->npml;
    const int pitch = ((MNAD_XO + g->nx + 2) + 31) / 32 * 32;
    const size_t elems = (size_t)(g->nz + 2 * MNAD_ZO + MNAD_ZPAD) * pitch;
    Shot s;
    base_desc(*g, pitch, nt, MODE_FWD, 2, s.d);
    std::vector<int> rrow, rcol;
    CHK(receiver_tables(*g, nr, zr, xr, s, rrow, rcol));
    std::vector<float> amp(nt + 1, 0.0f);
    {
        const double vsrc = v[(size_t)zs + (size_t)xs * g->nz];
        const double asrc = vsrc * vsrc * (g->dt * g->dt);   // fwd_PDE_exact_MNAD.m:39,111
        for (int t = 0; t < nt; ++t) {
            if (!std::isfinite(w[t])) return fail(MNAD_ERR_ARG, "wavelet sample %d is not finite", t);
            amp[t] = (float)(asrc * w[t]);
        }
    }

    // ---- scratch layout ----
    Carve cv;
    const size_t o_v = cv.take(n * sizeof(double));
    const size_t o_rx = cv.take(elems * 4);
    const size_t cfx_n = 3 * (size_t)g->nz * 2 * np, cfz_n = 3 * (size_t)2 * np * pitch;
    const size_t o_cfx = cv.take(cfx_n * 4), o_cfz = cv.take(cfz_n * 4);
    size_t o_f[6];
    for (int k = 0; k < 6; ++k) o_f[k] = cv.take(elems * 4);
    const size_t psx_n = 9 * (size_t)g->nz * 2 * np, psz_n = 9 * (size_t)2 * np * pitch;
    const size_t o_psx = cv.take(psx_n * 4), o_psz = cv.take(psz_n * 4);
    const size_t o_amp = cv.take(amp.size() * 4);
    const size_t o_rrow = cv.take(rrow.size() * 4), o_rcol = cv.take(rcol.size() * 4);
    const size_t o_gat = cv.take((size_t)nt * s.nu * 4);
    const size_t o_desc = cv.take(sizeof(ShotDev));
    const size_t o_base = cv.take(sizeof(int));
    // tile table last: its length is known once the descriptor is laid out
    auto dev = [&](size_t off) { return static_cast<void*>(base + off); };
    ShotDev& d = s.d;
    d.rx = static_cast<float*>(dev(o_rx));
    for (int k = 0; k < 6; ++k) d.f[k / 3][k % 3] = static_cast<float*>(dev(o_f[k]));
    if (np > 0) {
        d.psx = static_cast<float*>(dev(o_psx));
        d.psz = static_cast<float*>(dev(o_psz));
        d.cfx = static_cast<const float*>(dev(o_cfx));
        d.cfz = static_cast<const float*>(dev(o_cfz));
    }
    d.src_j = zs;
    d.src_i = xs;
    d.src_amp = static_cast<const float*>(dev(o_amp));
    if (s.nu > 0) {
        d.rrow = static_cast<const int*>(dev(o_rrow));
        d.rcol = static_cast<const int*>(dev(o_rcol));
        d.gather = static_cast<float*>(dev(o_gat));
    }
    std::vector<ShotDev> h{d};
    std::vector<TileRef> table;
    build_tiles(ctx, h, table);
    const size_t o_tab = cv.take(table.size() * sizeof(TileRef));
    if (cv.off > kScratchBytes)
        return fail(MNAD_ERR_ARG,
                    "the allocation-free forward of a %dx%d grid (npml %d), nt %d, %d receivers "
                    "needs %zu bytes of scratch; the array holds %zu (use mnad_fwd)",
                    g->nz, g->nx, np, nt, s.nu, cv.off, kScratchBytes);

    // ---- device work, all on the context stream ----
    hipStream_t st = ctx->stream;
    auto zero = [&](size_t off, size_t bytes) {
        return mnad_launch_fill(static_cast<float*>(dev(off)), (long)(bytes / 4), 0.0f, st);
    };
    HIPCHK(zero(o_rx, elems * 4));
    if (np > 0) {
        HIPCHK(zero(o_cfx, cfx_n * 4));
        HIPCHK(zero(o_cfz, cfz_n * 4));
        HIPCHK(zero(o_psx, psx_n * 4));
        HIPCHK(zero(o_psz, psz_n * 4));
    }
    for (int k = 0; k < 6; ++k) HIPCHK(zero(o_f[k], elems * 4));
    HIPCHK(zero(o_base, 4));   // step counter (0.0f has all bits clear)
    HIPCHK(hipMemcpyAsync(dev(o_v), v, n * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(mnad_launch_setup_model(static_cast<const double*>(dev(o_v)), g->nz, g->nx, pitch, np,
                                   g->dx, g->dz, g->dt, static_cast<float*>(dev(o_rx)), static_cast<float*>(dev(o_cfx)),
                                   static_cast<float*>(dev(o_cfz)), st));
    HIPCHK(hipMemcpyAsync(dev(o_amp), amp.data(), amp.size() * 4, hipMemcpyHostToDevice, st));
    if (s.nu > 0) {
        HIPCHK(hipMemcpyAsync(dev(o_rrow), rrow.data(), rrow.size() * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(dev(o_rcol), rcol.data(), rcol.size() * 4, hipMemcpyHostToDevice, st));
    }
    HIPCHK(hipMemcpyAsync(dev(o_desc), h.data(), sizeof(ShotDev), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dev(o_tab), table.data(), table.size() * sizeof(TileRef),
                          hipMemcpyHostToDevice, st));
    auto* ds = static_cast<ShotDev*>(dev(o_desc));
    HIPCHK(mnad_launch_init_source(ds, 1, st));
    const auto* tb = static_cast<const TileRef*>(dev(o_tab));
    const int* cnt = static_cast<const int*>(dev(o_base));
    for (int t = 0; t < nt; ++t)
        HIPCHK(mnad_launch_step(ds, tb, (int)table.size(), cnt, t, ctx->kernel, ctx->rz,
                                ctx->brows, st));

    // ---- outputs ----
    std::vector<float> gat((size_t)nt * s.nu);
    if (datau && s.nu > 0)
        HIPCHK(hipMemcpyAsync(gat.data(), dev(o_gat), gat.size() * 4, hipMemcpyDeviceToHost, st));
    double* outs[3] = {U, V, W};
    std::vector<float> fld[3];
    const int par = nt & 1;   // level nt lives in f[nt % 2]
    for (int k = 0; k < 3; ++k)
        if (outs[k]) {
            fld[k].resize(elems);
            HIPCHK(hipMemcpyAsync(fld[k].data(), dev(o_f[par * 3 + k]), elems * 4,
                                  hipMemcpyDeviceToHost, st));
        }
    HIPCHK(hipStreamSynchronize(st));   // also keeps the host sources alive until copied
    if (datau)
        for (int r = 0; r < nr; ++r) {
            const int u = s.uid[r];
            for (int t = 0; t < nt; ++t) datau[(size_t)t + (size_t)r * nt] = gat[(size_t)t * s.nu + u];
        }
    const double scale[3] = {1.0, 1.0 / g->dx, 1.0 / g->dz};   // Vs = dx V, Ws = dz W
    for (int k = 0; k < 3; ++k)
        if (outs[k])
            for (int i = 0; i < g->nx; ++i)
                for (int j = 0; j < g->nz; ++j)
                    outs[k][(size_t)j + (size_t)i * g->nz] =
                        scale[k] * fld[k][(size_t)(j + MNAD_ZO) * pitch + (i + MNAD_XO)];
    return MNAD_OK;
}
