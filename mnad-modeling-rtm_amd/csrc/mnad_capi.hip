// mnad_capi.hip -- the extern "C" entry points of include/mnad.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <vector>

#include "mnad.h"
#include "mnad_internal.h"

using namespace mnad;

namespace mnad {
extern "C" hipError_t mnad_launch_make_radj(const float* dobs, const float* dsyn,
                                            const float* arec, int nt, int nrec, float* radj,
                                            hipStream_t st);
extern "C" hipError_t mnad_launch_stack(const float* img, const float* imgc, const float* illum,
                                        float T, long n, float* isum, float* isumc, float* lsum,
                                        float* lsumc, hipStream_t st);
}

#define HIPCHK(x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess)                                                          \
            return fail(e_ == hipErrorOutOfMemory ? MNAD_ERR_NOMEM : MNAD_ERR_HIP,     \
                        "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                     \
    } while (0)

#define CHK(x)              \
    do {                    \
        int rc_ = (x);      \
        if (rc_ != MNAD_OK) \
            return rc_;     \
    } while (0)

#define ENTER(ctx)                                                   \
    if (!(ctx)) return fail(MNAD_ERR_ARG, "context is NULL");        \
    std::lock_guard<std::recursive_mutex> lock_((ctx)->mu);          \
    HIPCHK(hipSetDevice((ctx)->device))

static bool tuning_ok(int kernel, int rows, int brows)
{
    if (kernel == 0 && brows == 4) return rows == 1 || rows == 2 || rows == 4;
    if (kernel == 0 && (brows == 1 || brows == 2 || brows == 8 || brows == 16)) return rows == 1;
    if (kernel == 4) return rows == 1 && (brows == 1 || brows == 2 || brows == 4 || brows == 8);
    if (kernel == 6 || kernel == 10) return rows == 1 && (brows == 1 || brows == 2 || brows == 4);
    if (kernel == 7)   // output rows 4 * rows; (rows + 1) * 4 phase-1 rows split over block_rows waves
        return (rows == 3 && (brows == 4 || brows == 8 || brows == 16)) ||
               (rows == 5 && (brows == 4 || brows == 8)) ||
               (rows == 7 && (brows == 8 || brows == 16));
    if (kernel == 9) return brows == 8 && rows >= 1 && rows <= 4;   // rows: blocks per CU
    if (kernel == 11) return brows == 1 && rows == 2;                // two rows per wavefront
    if (kernel == 13)   // register-queue two-step strips: segments of 8 * rows rows
        return (brows == 1 || brows == 2 || brows == 4) && rows >= 1 && rows <= 256;
    return false;
}

extern "C" {

const char* mnad_version(void) { return "libmnad 0.1.0 (gfx950)"; }

const char* mnad_last_error(void) { return last_error(); }

int mnad_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int mnad_ctx_create(int device, mnad_ctx** out)
{
    if (!out) return fail(MNAD_ERR_ARG, "out is NULL");
    *out = nullptr;
    int n = mnad_device_count();
    if (n <= 0) return fail(MNAD_ERR_NOGPU, "no HIP device visible: libmnad needs a GPU");
    if (device < 0 || device >= n)
        return fail(MNAD_ERR_ARG, "device %d out of range (%d visible)", device, n);
    HIPCHK(hipSetDevice(device));
    mnad_ctx* c = new mnad_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(MNAD_ERR_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    if (int rc = noalloc_init_device(device); rc != MNAD_OK) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    if (const char* s = getenv("MNAD_GRAPH_CHUNK")) c->graph_chunk = std::max(1, atoi(s));
    if (const char* s = getenv("MNAD_KERNEL")) c->kernel = atoi(s);
    if (const char* s = getenv("MNAD_RZ")) c->rz = atoi(s);
    if (const char* s = getenv("MNAD_BROWS")) c->brows = atoi(s);
    if (!tuning_ok(c->kernel, c->rz, c->brows)) {
        c->kernel = kDefaultKernel;
        c->rz = kDefaultRows;
        c->brows = kDefaultBrows;
    }
    if (const char* s = getenv("MNAD_STRIP_FIRST")) c->strip_first = atoi(s) != 0;
    if (const char* s = getenv("MNAD_STRIP_WEIGHT")) c->strip_weight = std::max(1.0, atof(s));
    if (const char* s = getenv("MNAD_X4_WEIGHT")) c->x4_weight = std::max(1.0, atof(s));
    if (const char* s = getenv("MNAD_AUTOTUNE")) c->autotune_on_use = atoi(s) != 0;
    if (const char* s = getenv("MNAD_RING_ORDER")) c->ring_order = std::max(2, std::min(4, atoi(s)));
    *out = c;
    return MNAD_OK;
}

void mnad_ctx_destroy(mnad_ctx* ctx)
{
    if (!ctx) return;
    server_destroy(ctx);
    {
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        ctx->kept.clear();
        ctx->stacks.clear();
        ctx->models.clear();
        release_batch_cache(ctx);
        ctx->pool.trim();
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

int mnad_ctx_sync(mnad_ctx* ctx)
{
    ENTER(ctx);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MNAD_OK;
}

int mnad_snap_schedule(int nt, double vmin, double vmax, int* nf, int* T, int* K, int* ncol)
{
    if (!(vmin > 0) || !(vmax >= vmin)) return fail(MNAD_ERR_ARG, "bad velocity range");
    return schedule(nt, vmin, vmax, nf, T, K, ncol);
}

int mnad_timing(mnad_ctx* ctx, int reset, double* ms, int64_t* launches, double* bytes)
{
    ENTER(ctx);
    if (ms) *ms = ctx->t_ms;
    if (launches) *launches = ctx->t_launches;
    if (bytes) *bytes = ctx->t_bytes;
    if (reset) {
        ctx->t_ms = 0.0;
        ctx->t_launches = 0;
        ctx->t_bytes = 0.0;
    }
    return MNAD_OK;
}

int mnad_get_tuning(mnad_ctx* ctx, mnad_tuning* t)
{
    ENTER(ctx);
    if (!t) return fail(MNAD_ERR_ARG, "tuning is NULL");
    t->kernel = ctx->kernel;
    t->rows_per_thread = ctx->rz;
    t->block_rows = ctx->brows;
    t->graph_chunk = ctx->graph_chunk;
    t->strip_first = ctx->strip_first;
    return MNAD_OK;
}

int mnad_set_tuning(mnad_ctx* ctx, const mnad_tuning* t)
{
    ENTER(ctx);
    if (!t) return fail(MNAD_ERR_ARG, "tuning is NULL");
    if (!tuning_ok(t->kernel, t->rows_per_thread, t->block_rows))
        return fail(MNAD_ERR_ARG,
                    "unsupported step kernel %d with %d rows per thread and %d block rows",
                    t->kernel, t->rows_per_thread, t->block_rows);
    ctx->kernel = t->kernel;
    ctx->rz = t->rows_per_thread;
    ctx->brows = t->block_rows;
    ctx->graph_chunk = std::max(1, t->graph_chunk);
    ctx->strip_first = t->strip_first != 0;
    return MNAD_OK;
}

int mnad_mem_info(mnad_ctx* ctx, uint64_t* in_use, uint64_t* cached)
{
    if (!ctx) return fail(MNAD_ERR_ARG, "context is NULL");
    std::lock_guard<std::mutex> lk(ctx->pool.mu);
    if (in_use) *in_use = ctx->pool.in_use;
    if (cached) *cached = ctx->pool.cached;
    return MNAD_OK;
}

// fwd_PDE_exact_MNAD.m:1
int mnad_fwd(mnad_ctx* ctx, const mnad_grid* g, const double* v, int nt, const double* w,
             int zs, int xs, int nr, const int* zr, const int* xr, int ss, int indx,
             int snap_fmt, int flags, double* datau, double* U, double* V, double* W,
             double* illum)
{
    ENTER(ctx);
    if (indx != 0 && indx != 1) return fail(MNAD_ERR_ARG, "indx must be 0 or 1");
    if (!snap_format_ok(snap_fmt)) return fail(MNAD_ERR_ARG, "unknown snapshot format %d", snap_fmt);
    if (nr < 0 || (nr > 0 && (!zr || !xr))) return fail(MNAD_ERR_ARG, "bad receiver arrays");
    Model m;
    CHK(make_model(ctx, g, v, m));
    ShotSpec sp;
    sp.mode = MODE_FWD;
    sp.nt = nt;
    sp.w = w;
    sp.zs = zs;
    sp.xs = xs;
    sp.nr = nr;
    sp.zr = zr;
    sp.xr = xr;
    const bool ring = indx == 1 && ring_format(snap_fmt);
    if (indx == 1) {
        CHK(schedule(nt, m.vmin, m.vmax, nullptr, &sp.T, &sp.K, &sp.ncol));
        sp.snap_fmt = snap_fmt;
        sp.ring_order = ctx->ring_order;
    }
    // the ring code runs on the one-step kernels 4, 5, 6, 10, 11
    TunePick ring_pick{kDefaultKernel, kDefaultRows, kDefaultBrows, 1, 0.0, {}};
    TuningGuard tg(ctx, ring && !ring_kernel(ctx->kernel) ? &ring_pick : nullptr);
    auto sp_shot = std::make_shared<Shot>();
    Shot& s = *sp_shot;
    CHK(make_shot(ctx, m, sp, s));
    std::vector<Shot*> b{&s};
    CHK(run_batch(ctx, b, flags));
    CHK(download_gather(ctx, s, nt, nr, datau));
    CHK(download_field(ctx, m, s, 0, 1.0, U));
    CHK(download_field(ctx, m, s, 1, 1.0 / g->dx, V));
    CHK(download_field(ctx, m, s, 2, 1.0 / g->dz, W));
    if (indx == 1) {
        if (illum) CHK(download_interior(ctx, *g, s.illum.as<float>(), nullptr, (double)sp.T, illum));
        Kept k;
        k.snaps = std::move(s.snaps);
        k.fmt = snap_fmt;
        k.sscale = s.d.sscale;
        k.nz = g->nz;
        k.nx = g->nx;
        k.npml = g->npml;
        k.nt = nt;
        k.T = sp.T;
        k.K = sp.K;
        k.ncol = sp.ncol;
        k.dx = g->dx;
        k.dz = g->dz;
        k.dt = g->dt;
        if (ring) {   // keep the forward's last two levels, source and ring only
            s.psx.reset();
            s.psz.reset();
            s.gather.reset();
            s.rrow.reset();
            s.rcol.reset();
            s.illum.reset();
            // (the field planes are one allocation: the spare levels of a
            // two-step kernel's shot stay with it)
            k.fwd = sp_shot;
        }
        ctx->kept[ss] = std::move(k);
    }
    return MNAD_OK;
}

// Born_adj_MNAD.m:1 (U_r := U)
int mnad_born_adj(mnad_ctx* ctx, const mnad_grid* g, const double* v, int nt, int nr,
                  const int* zr, const int* xr, const double* data, int ss, int flags, double* Im)
{
    ENTER(ctx);
    if (nr < 0 || (nr > 0 && (!zr || !xr || !data))) return fail(MNAD_ERR_ARG, "bad receiver arrays");
    auto it = ctx->kept.find(ss);
    if (it == ctx->kept.end())
        return fail(MNAD_ERR_STATE,
                    "no source wavefield for shot %d: run mnad_fwd(..., ss=%d, indx=1) first "
                    "(Born_adj_MNAD.m:33 loads ./Wavefields/UU_<ss>)",
                    ss, ss);
    Model m;
    CHK(make_model(ctx, g, v, m));
    ShotSpec sp;
    sp.mode = MODE_ADJ;
    sp.nt = nt;
    sp.nr = nr;
    sp.zr = zr;
    sp.xr = xr;
    CHK(schedule(nt, m.vmin, m.vmax, nullptr, &sp.T, &sp.K, &sp.ncol));
    Kept& k = it->second;
    if (k.nz != g->nz || k.nx != g->nx || k.npml != g->npml || k.nt != nt || k.T != sp.T ||
        k.K != sp.K || k.ncol != sp.ncol)
        return fail(MNAD_ERR_STATE,
                    "shot %d: kept snapshots (grid %dx%d, npml %d, nt %d, T %d, K %d) do not "
                    "match this call (grid %dx%d, npml %d, nt %d, T %d, K %d)",
                    ss, k.nz, k.nx, k.npml, k.nt, k.T, k.K, g->nz, g->nx, g->npml, nt, sp.T, sp.K);
    if (k.fwd) {   // ring storage: reverse reconstruction beside the adjoint
        TunePick ring_pick{kDefaultKernel, kDefaultRows, kDefaultBrows, 1, 0.0, {}};
        // the kept forward's levels (two, or four for a two-step kernel) fix
        // the kernel family the reconstruction runs on
        const bool fits = ring_kernel(ctx->kernel) &&
                          k.fwd->d.nlev == (two_step_kernel(ctx->kernel) ? 4 : 2);
        TuningGuard tg(ctx, fits ? nullptr : &ring_pick);
        Shot s;
        CHK(make_shot(ctx, m, sp, s));
        Shot& f = *k.fwd;
        f.d.rx = m.rx();   // the migration model of this call (the forward's, by contract)
        // The reconstruction runs the kept levels and ring state backwards; a
        // copy taken first restores them if the adjoint fails, so the source
        // wavefield survives for a retry as the reference's UU_<ss>.mat does.
        DevBuf bk[7];
        for (int q = 0; q < 6; ++q) {
            CHK(dalloc(ctx, bk[q], f.state[q].n, false));
            HIPCHK(hipMemcpyAsync(bk[q].p, f.state[q].p, f.state[q].n, hipMemcpyDeviceToDevice,
                                  ctx->stream));
        }
        CHK(dalloc(ctx, bk[6], f.rst.n, false));
        HIPCHK(hipMemcpyAsync(bk[6].p, f.rst.p, f.rst.n, hipMemcpyDeviceToDevice, ctx->stream));
        Shot rec;
        int rc = make_rec_shot(ctx, f, s, rec);
        if (rc == MNAD_OK) rc = fill_radj_host(ctx, s, nt, nr, data);
        if (rc == MNAD_OK) rc = run_batch(ctx, std::vector<Shot*>{&s, &rec}, flags);
        if (rc == MNAD_OK)
            rc = download_interior(ctx, *g, s.img.as<float>(), s.imgc.as<float>(), (double)sp.T, Im);
        if (rc != MNAD_OK) {
            (void)hipStreamSynchronize(ctx->stream);
            for (int q = 0; q < 6; ++q)
                if (rec.state[q].p) f.state[q] = std::move(rec.state[q]);
            if (rec.amp.p) f.amp = std::move(rec.amp);
            if (rec.rq.p) f.rq = std::move(rec.rq);
            if (rec.rst.p) f.rst = std::move(rec.rst);
            f.d.rq = f.rq.as<unsigned short>();
            f.d.rst = f.rst.as<long long>();
            for (int q = 0; q < 6; ++q)
                (void)hipMemcpyAsync(f.state[q].p, bk[q].p, f.state[q].n, hipMemcpyDeviceToDevice,
                                     ctx->stream);
            (void)hipMemcpyAsync(f.rst.p, bk[6].p, f.rst.n, hipMemcpyDeviceToDevice, ctx->stream);
            (void)hipStreamSynchronize(ctx->stream);
            return rc;
        }
        ctx->kept.erase(it);
        return MNAD_OK;
    }
    Shot s;
    CHK(make_shot(ctx, m, sp, s));
    // The kept source wavefield is lent to the adjoint and released only once
    // the image is out: a failed adjoint leaves it in place for a retry, as the
    // reference's UU_<ss>.mat survives a failed Born_adj_MNAD.
    attach_snaps(s, std::move(k.snaps), k.fmt, k.sscale);
    int rc = fill_radj_host(ctx, s, nt, nr, data);
    if (rc == MNAD_OK) rc = run_batch(ctx, std::vector<Shot*>{&s}, flags);
    if (rc == MNAD_OK)
        rc = download_interior(ctx, *g, s.img.as<float>(), s.imgc.as<float>(), (double)sp.T, Im);
    if (rc != MNAD_OK) {
        (void)hipStreamSynchronize(ctx->stream);   // nothing queued still reads it
        k.snaps = std::move(s.snaps);
        s.d.snaps = nullptr;
        return rc;
    }
    ctx->kept.erase(it);
    return MNAD_OK;
}

int mnad_release(mnad_ctx* ctx, int ss)
{
    ENTER(ctx);
    ctx->kept.erase(ss);
    return MNAD_OK;
}

namespace {
struct SnapHeader {
    char magic[8];   // "MNADSNP1"
    int32_t nz, nx, npml, nt, T, K, ncol, fmt;
    float sscale;
    int32_t pad;
    uint64_t bytes;
};

// Ring storage: the forward's last two levels (6 field planes), source
// amplitudes, ring codes and predictor state.
struct RingHeader {
    char magic[8];   // "MNADRNG1"
    int32_t nz, nx, npml, nt, T, K, ncol, fmt;
    int32_t rord, nring, rfmt, src_j, src_i, pitch, pad0, pad1;
    double dx, dz, dt, rsc, rcs;
    uint64_t plane_bytes, amp_bytes, rq_bytes, rst_bytes;
};

// Device <-> file through a bounded host buffer, on the context stream.
int write_dev(mnad_ctx* ctx, FILE* f, const void* p, size_t bytes)
{
    std::vector<char> buf(std::min<size_t>(bytes, (size_t)256 << 20));
    for (size_t off = 0; off < bytes; off += buf.size()) {
        const size_t n = std::min(buf.size(), bytes - off);
        if (hipMemcpyAsync(buf.data(), static_cast<const char*>(p) + off, n, hipMemcpyDeviceToHost,
                           ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            return fail(MNAD_ERR_HIP, "copy of the source wavefield to the host failed");
        if (fwrite(buf.data(), 1, n, f) != n) return fail(MNAD_ERR_ARG, "file write failed");
    }
    return MNAD_OK;
}

int read_dev(mnad_ctx* ctx, FILE* f, void* p, size_t bytes, const char* path)
{
    std::vector<char> buf(std::min<size_t>(bytes, (size_t)256 << 20));
    for (size_t off = 0; off < bytes; off += buf.size()) {
        const size_t n = std::min(buf.size(), bytes - off);
        if (fread(buf.data(), 1, n, f) != n) return fail(MNAD_ERR_ARG, "%s is truncated", path);
        if (hipMemcpyAsync(static_cast<char*>(p) + off, buf.data(), n, hipMemcpyHostToDevice,
                           ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            return fail(MNAD_ERR_HIP, "copy of the source wavefield to the device failed");
    }
    return MNAD_OK;
}

int ring_save(mnad_ctx* ctx, const Kept& k, const char* path)
{
    const Shot& s = *k.fwd;
    RingHeader h{};
    memcpy(h.magic, "MNADRNG1", 8);
    h.nz = k.nz; h.nx = k.nx; h.npml = k.npml; h.nt = k.nt; h.T = k.T; h.K = k.K;
    h.ncol = k.ncol; h.fmt = k.fmt;
    h.rord = s.d.rord; h.nring = s.d.nring; h.rfmt = s.d.rfmt;
    h.src_j = s.d.src_j; h.src_i = s.d.src_i; h.pitch = s.d.pitch;
    h.dx = k.dx; h.dz = k.dz; h.dt = k.dt; h.rsc = s.d.rsc; h.rcs = s.d.rcs;
    h.plane_bytes = plane_stride_bytes(k.nz, s.d.pitch);
    h.amp_bytes = s.amp.n; h.rq_bytes = s.rq.n; h.rst_bytes = s.rst.n;
    FILE* f = fopen(path, "wb");
    if (!f) return fail(MNAD_ERR_ARG, "cannot open %s for writing", path);
    int rc = fwrite(&h, sizeof(h), 1, f) == 1 ? MNAD_OK : fail(MNAD_ERR_ARG, "write to %s failed", path);
    // the forward's last two levels, nt - 1 and nt, as a two-level shot keeps
    // them (level L in slot L mod 2), U, Vs, Ws each, one plane after the other
    for (int q = 0; q < 6 && rc == MNAD_OK; ++q) {
        const int L = ((k.nt - 1) % 2 == q / 3) ? k.nt - 1 : k.nt;
        rc = write_dev(ctx, f, s.d.f[L % s.d.nlev][q % 3], h.plane_bytes);
    }
    if (rc == MNAD_OK) rc = write_dev(ctx, f, s.amp.p, s.amp.n);
    if (rc == MNAD_OK) rc = write_dev(ctx, f, s.rq.p, s.rq.n);
    if (rc == MNAD_OK) rc = write_dev(ctx, f, s.rst.p, s.rst.n);
    if (fclose(f) != 0 && rc == MNAD_OK) rc = fail(MNAD_ERR_ARG, "write to %s failed", path);
    return rc;
}

int ring_load(mnad_ctx* ctx, FILE* f, const char* path, Kept& k)
{
    RingHeader h{};
    memcpy(h.magic, "MNADRNG1", 8);
    if (fread((char*)&h + 8, sizeof(h) - 8, 1, f) != 1)
        return fail(MNAD_ERR_ARG, "%s is not a snapshot file", path);
    mnad_grid g{h.nz, h.nx, h.npml, h.dx, h.dz, h.dt};
    const int pitch = ((MNAD_XO + h.nx + 2) + 31) / 32 * 32;
    const size_t plane = plane_stride_bytes(h.nz, pitch);
    const size_t nr = (h.npml >= 2 && h.nz > 2 * h.npml && h.nx > 2 * h.npml)
                          ? (size_t)ring_cells(h.nz, h.nx, h.npml) : 0;
    auto r256 = [](size_t b) { return (b + 255) & ~(size_t)255; };
    if (!ring_format(h.fmt) || h.nt < 1 || h.rord < 2 || h.rord > RING_MAX_ORDER || nr == 0 ||
        (int)nr != h.nring || h.pitch != pitch || h.plane_bytes != plane ||
        h.amp_bytes != r256((size_t)(h.nt + 1) * 4) || h.rq_bytes != r256((size_t)h.nt * 3 * nr * 2) ||
        h.rst_bytes != r256((size_t)h.rord * 3 * nr * 8) || h.T < 1 || h.K < 1 ||
        !(h.rsc > 0) || !(h.rcs > 0))
        return fail(MNAD_ERR_ARG, "%s: inconsistent header", path);
    auto s = std::make_shared<Shot>();
    base_desc(g, pitch, h.nt, MODE_FWD, 2, s->d);
    ShotDev& d = s->d;
    CHK(dalloc(ctx, s->state[0], 6 * plane, false));   // one allocation, as make_shot's
    d.pbytes = (unsigned)plane;
    for (int q = 0; q < 6; ++q) {
        d.f[q / 3][q % 3] = s->state[0].as<float>() + q * (plane / sizeof(float));
        CHK(read_dev(ctx, f, d.f[q / 3][q % 3], plane, path));
    }
    CHK(dalloc(ctx, s->amp, h.amp_bytes, false));
    CHK(read_dev(ctx, f, s->amp.p, h.amp_bytes, path));
    CHK(dalloc(ctx, s->rq, h.rq_bytes, false));
    CHK(read_dev(ctx, f, s->rq.p, h.rq_bytes, path));
    CHK(dalloc(ctx, s->rst, h.rst_bytes, false));
    CHK(read_dev(ctx, f, s->rst.p, h.rst_bytes, path));
    d.src_j = h.src_j;
    d.src_i = h.src_i;
    d.src_amp = s->amp.as<float>();
    d.T = h.T;
    d.K = h.K;
    d.shift = h.ncol - h.K;
    d.rord = h.rord;
    d.nring = h.nring;
    d.rfmt = h.rfmt;
    d.rq = s->rq.as<unsigned short>();
    d.rst = s->rst.as<long long>();
    d.rsc = h.rsc;
    d.rsi = 1.0 / h.rsc;
    d.rcs = h.rcs;
    d.rci = 1.0 / h.rcs;
    k.fmt = h.fmt; k.nz = h.nz; k.nx = h.nx; k.npml = h.npml; k.nt = h.nt; k.T = h.T;
    k.K = h.K; k.ncol = h.ncol; k.dx = h.dx; k.dz = h.dz; k.dt = h.dt;
    k.fwd = s;
    return MNAD_OK;
}
}  // namespace

int mnad_snap_save(mnad_ctx* ctx, int ss, const char* path)
{
    ENTER(ctx);
    if (!path) return fail(MNAD_ERR_ARG, "path is NULL");
    auto it = ctx->kept.find(ss);
    if (it == ctx->kept.end()) return fail(MNAD_ERR_STATE, "no snapshots kept for shot %d", ss);
    const Kept& k = it->second;
    if (k.fwd) return ring_save(ctx, k, path);
    const size_t ni = (size_t)(k.nz - 2 * k.npml) * (k.nx - 2 * k.npml);
    const size_t bytes = (size_t)k.K * ni * (k.fmt == MNAD_SNAP_F32 ? 4 : 2);
    SnapHeader h{};
    memcpy(h.magic, "MNADSNP1", 8);
    h.nz = k.nz; h.nx = k.nx; h.npml = k.npml; h.nt = k.nt; h.T = k.T; h.K = k.K;
    h.ncol = k.ncol; h.fmt = k.fmt; h.sscale = k.sscale; h.bytes = bytes;
    FILE* f = fopen(path, "wb");
    if (!f) return fail(MNAD_ERR_ARG, "cannot open %s for writing", path);
    bool ok = fwrite(&h, sizeof(h), 1, f) == 1;
    std::vector<char> buf(std::min<size_t>(bytes, (size_t)256 << 20));
    for (size_t off = 0; ok && off < bytes; off += buf.size()) {
        const size_t n = std::min(buf.size(), bytes - off);
        // on the context stream, behind every kernel that may still write the block
        if (hipMemcpyAsync(buf.data(), static_cast<const char*>(k.snaps.p) + off, n,
                           hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess) {
            fclose(f);
            return fail(MNAD_ERR_HIP, "copy of snapshots to the host failed");
        }
        ok = fwrite(buf.data(), 1, n, f) == n;
    }
    ok = (fclose(f) == 0) && ok;
    if (!ok) return fail(MNAD_ERR_ARG, "write to %s failed", path);
    return MNAD_OK;
}

int mnad_snap_load(mnad_ctx* ctx, int ss, const char* path)
{
    ENTER(ctx);
    if (!path) return fail(MNAD_ERR_ARG, "path is NULL");
    FILE* f = fopen(path, "rb");
    if (!f) return fail(MNAD_ERR_STATE, "cannot open %s", path);
    char magic[8];
    if (fread(magic, 8, 1, f) != 1) {
        fclose(f);
        return fail(MNAD_ERR_ARG, "%s is not a snapshot file", path);
    }
    if (memcmp(magic, "MNADRNG1", 8) == 0) {
        Kept k;
        const int rc = ring_load(ctx, f, path, k);
        fclose(f);
        if (rc != MNAD_OK) return rc;
        ctx->kept[ss] = std::move(k);
        return MNAD_OK;
    }
    SnapHeader h{};
    memcpy(h.magic, magic, 8);
    if (fread((char*)&h + 8, sizeof(h) - 8, 1, f) != 1 || memcmp(h.magic, "MNADSNP1", 8) != 0 || h.K < 0 ||
        h.fmt < MNAD_SNAP_F32 || h.fmt > MNAD_SNAP_BF16 || h.nz <= 2 * h.npml || h.nx <= 2 * h.npml) {
        fclose(f);
        return fail(MNAD_ERR_ARG, "%s is not a snapshot file", path);
    }
    const size_t ni = (size_t)(h.nz - 2 * h.npml) * (h.nx - 2 * h.npml);
    const size_t bytes = (size_t)h.K * ni * (h.fmt == MNAD_SNAP_F32 ? 4 : 2);
    if (bytes != h.bytes) {
        fclose(f);
        return fail(MNAD_ERR_ARG, "%s: inconsistent header", path);
    }
    Kept k;
    int rc = dalloc(ctx, k.snaps, bytes, false);
    if (rc != MNAD_OK) {
        fclose(f);
        return rc;
    }
    std::vector<char> buf(std::min<size_t>(bytes, (size_t)256 << 20));
    for (size_t off = 0; off < bytes; off += buf.size()) {
        const size_t n = std::min(buf.size(), bytes - off);
        if (fread(buf.data(), 1, n, f) != n) {
            fclose(f);
            return fail(MNAD_ERR_ARG, "%s is truncated", path);
        }
        // on the context stream: the pool hands blocks back in stream order, so a
        // recycled block may still be read by a kernel queued there
        if (hipMemcpyAsync(static_cast<char*>(k.snaps.p) + off, buf.data(), n,
                           hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess) {
            fclose(f);
            return fail(MNAD_ERR_HIP, "copy of snapshots to the device failed");
        }
    }
    fclose(f);
    k.fmt = h.fmt; k.sscale = h.sscale; k.nz = h.nz; k.nx = h.nx; k.npml = h.npml;
    k.nt = h.nt; k.T = h.T; k.K = h.K; k.ncol = h.ncol;
    ctx->kept[ss] = std::move(k);
    return MNAD_OK;
}

int mnad_stack_create(mnad_ctx* ctx, const mnad_grid* g, int* stack_id)
{
    ENTER(ctx);
    if (!g || !stack_id) return fail(MNAD_ERR_ARG, "NULL argument");
    Stack st;
    st.g = *g;
    const size_t ni = (size_t)(g->nz - 2 * g->npml) * (g->nx - 2 * g->npml);
    if (g->nz <= 2 * g->npml || g->nx <= 2 * g->npml) return fail(MNAD_ERR_ARG, "empty interior");
    CHK(dalloc(ctx, st.isum, ni * sizeof(float), true));
    CHK(dalloc(ctx, st.isumc, ni * sizeof(float), true));
    CHK(dalloc(ctx, st.lsum, ni * sizeof(float), true));
    CHK(dalloc(ctx, st.lsumc, ni * sizeof(float), true));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const int id = ctx->next_stack++;
    ctx->stacks.emplace(id, std::move(st));
    *stack_id = id;
    return MNAD_OK;
}

int mnad_stack_destroy(mnad_ctx* ctx, int stack_id)
{
    ENTER(ctx);
    ctx->stacks.erase(stack_id);
    return MNAD_OK;
}

// main_Acoustic_RTM.m:243,256-262: img = mean_s(Im_s) ./ mean_s(illum_s).
int mnad_stack_image(mnad_ctx* ctx, int stack_id, double* img, double* illum, int* nshots)
{
    ENTER(ctx);
    auto it = ctx->stacks.find(stack_id);
    if (it == ctx->stacks.end()) return fail(MNAD_ERR_STATE, "unknown stack %d", stack_id);
    Stack& st = it->second;
    const mnad_grid& g = st.g;
    const size_t ni = (size_t)(g.nz - 2 * g.npml) * (g.nx - 2 * g.npml);
    std::vector<double> isum(ni), lsum(ni);
    CHK(download_interior(ctx, g, st.isum.as<float>(), st.isumc.as<float>(), 1.0, isum.data()));
    CHK(download_interior(ctx, g, st.lsum.as<float>(), st.lsumc.as<float>(), 1.0, lsum.data()));
    const double ns = st.nshots > 0 ? st.nshots : 1;
    for (size_t q = 0; q < ni; ++q) {
        if (img) img[q] = (isum[q] / ns) / (lsum[q] / ns);
        if (illum) illum[q] = lsum[q] / ns;
    }
    if (nshots) *nshots = st.nshots;
    return MNAD_OK;
}

}  // extern "C"

// ------------------------------------------------------------- batched RTM
namespace {

size_t shot_bytes(const Model& m, int nt, int nu)
{
    const mnad_grid& g = m.g;
    const size_t np = g.npml;
    return 6 * m.elems * 4 + 9 * (g.nz * 2 * np + 2 * np * g.nx) * 4 + (size_t)nt * nu * 4 + 4096;
}

// Shots whose per-shot bytes fit the context's budget (at least 1).
int mem_cap(mnad_ctx* ctx, size_t per_shot)
{
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) {
        (void)hipGetLastError();
        freeb = 0;
    }
    size_t budget = ctx->budget ? ctx->budget : (size_t)(0.85 * (double)(freeb + ctx->pool.cached));
    return (int)std::min<size_t>(1 << 20, std::max<size_t>(1, budget / std::max<size_t>(per_shot, 1)));
}

int pick_batch(mnad_ctx* ctx, size_t per_shot, int nshots, int batch)
{
    int b = mem_cap(ctx, per_shot);
    if (batch > 0) b = std::min(b, batch);
    b = std::max(1, std::min(b, nshots));
    const int groups = (nshots + b - 1) / b;     // even groups: no small last batch
    return (nshots + groups - 1) / groups;
}

int check_shots(const mnad_shots* sh)
{
    if (!sh || sh->nshots < 1 || !sh->zs || !sh->xs || !sh->rec_off)
        return fail(MNAD_ERR_ARG, "bad shot list");
    if (sh->rec_off[0] != 0) return fail(MNAD_ERR_ARG, "rec_off[0] must be 0");
    for (int k = 0; k < sh->nshots; ++k)
        if (sh->rec_off[k + 1] < sh->rec_off[k]) return fail(MNAD_ERR_ARG, "rec_off not monotone");
    if (sh->rec_off[sh->nshots] > 0 && (!sh->zr || !sh->xr))
        return fail(MNAD_ERR_ARG, "receiver arrays are NULL");
    return MNAD_OK;
}

}  // namespace

extern "C" int mnad_rtm_shots(mnad_ctx* ctx, int stack_id, const mnad_grid* g,
                              const double* v_true, const double* v_smooth, int nt,
                              const double* w, const mnad_shots* sh, int snap_fmt, int batch,
                              int flags, double* dobs, double* dsyn)
{
    ENTER(ctx);
    CHK(check_shots(sh));
    if (!snap_format_ok(snap_fmt)) return fail(MNAD_ERR_ARG, "unknown snapshot format %d", snap_fmt);
    const bool ring = ring_format(snap_fmt);
    if (ring && g && g->npml < 2)
        return fail(MNAD_ERR_ARG, "ring storage needs npml >= 2 (npml = %d)", g->npml);
    auto sit = ctx->stacks.find(stack_id);
    if (sit == ctx->stacks.end()) return fail(MNAD_ERR_STATE, "unknown stack %d", stack_id);
    Stack& st = sit->second;
    if (st.g.nz != g->nz || st.g.nx != g->nx || st.g.npml != g->npml)
        return fail(MNAD_ERR_ARG, "stack grid does not match");
    Model mt, ms;
    CHK(make_model(ctx, g, v_true, mt));
    CHK(make_model(ctx, g, v_smooth, ms));
    int T, K, ncol;
    CHK(schedule(nt, ms.vmin, ms.vmax, nullptr, &T, &K, &ncol));   // from vs, as :224
    const size_t ni = (size_t)(g->nz - 2 * g->npml) * (g->nx - 2 * g->npml);
    int maxnr = 0;
    for (int k = 0; k < sh->nshots; ++k) maxnr = std::max(maxnr, sh->rec_off[k + 1] - sh->rec_off[k]);
    const size_t esz = snap_fmt == MNAD_SNAP_F32 ? 4 : 2;
    const bool serial = (flags & MNAD_FLAG_SERIAL) != 0;
    // Two schedules for groups of B shots.  In order: one launch sequence
    // runs the group's 2B forwards, the next its B adjoints.  Pipelined: one
    // sequence runs the group's forwards together with the previous group's
    // adjoints (3B propagations), so two groups' snapshots are alive at once.
    // Ring storage: the adjoints' launches also run the B reverse
    // reconstructions, which keep the forwards' fields, ring and a slot.
    const size_t keep = ring ? (size_t)nt * 3 * ring_cells(g->nz, g->nx, g->npml) * 2 +
                                   (size_t)ctx->ring_order * 3 * ring_cells(g->nz, g->nx, g->npml) * 8 +
                                   ni * 4
                             : (size_t)K * ni * esz;
    const size_t per_ord = 3 * shot_bytes(mt, nt, maxnr) + keep + 4 * ni * 4;
    const size_t per_pip = per_ord + keep;
    const TunePick* tp = nullptr;
    if (!serial) {
        tp = find_tuned(ctx, *g);
        if (!tp && batch == 0 && ((flags & MNAD_FLAG_AUTOTUNE) || ctx->autotune_on_use)) {
            TunePick pick;
            CHK(autotune(ctx, ms, 32, &pick));
            tp = find_tuned(ctx, *g);
        }
    }
    // the ring code and the reverse reconstruction run on the one-step kernels
    // 4, 5, 6, 10, 11; a tuned pick of another family gives way to kernel 4
    TunePick ring_pick{kDefaultKernel, kDefaultRows, kDefaultBrows, 1, 0.0, {}};
    if (ring && tp && !ring_kernel(tp->kernel)) tp = nullptr;
    if (ring && !tp && !ring_kernel(ctx->kernel)) tp = &ring_pick;
    TuningGuard tg(ctx, tp);
    if (tp == &ring_pick) tp = nullptr;   // no time-per-batch curve for the schedule
    bool pipelined = !serial;
    if (!serial && batch == 0 && tp) {
        // the tuned kernel's time per shot-step against shots per launch:
        // per shot, in order costs 2 t(2B) + t(B), pipelined 3 t(3B)
        double best = 1e300;
        const int cap_ord = mem_cap(ctx, per_ord), cap_pip = mem_cap(ctx, per_pip);
        for (int b = 1; b <= sh->nshots; ++b) {
            const double t2 = tp->at(2 * b), t1 = tp->at(b), t3 = tp->at(3 * b);
            if (b <= cap_ord && t2 > 0 && t1 > 0 && 2 * t2 + t1 < best) {
                best = 2 * t2 + t1;
                batch = b;
                pipelined = false;
            }
            if (b <= cap_pip && t3 > 0 && 3 * t3 <= best) {
                best = 3 * t3;
                batch = b;
                pipelined = true;
            }
        }
        if (batch == 0) batch = 1;
    }
    if (serial) batch = 1;
    const int B = pick_batch(ctx, pipelined ? per_pip : per_ord, sh->nshots, batch);
    ctx->last_group = B;
    ctx->last_pipelined = pipelined ? 1 : 0;
    if (getenv("MNAD_AUTOTUNE_LOG"))
        fprintf(stderr, "mnad_rtm_shots %dx%d: %d shots, %s, %d per group\n", g->nz, g->nx,
                sh->nshots, serial ? "serial" : (pipelined ? "pipelined" : "in order"), B);

    struct Group {
        int k0 = 0, nb = 0;
        std::vector<Shot> fw, ad;   // fw[2b]: observed (true model), fw[2b+1]: background
        std::vector<Shot> rc;       // ring storage: reverse reconstructions of fw[2b+1]
    };
    // forward shots of shots [k0, k0 + nb)
    auto forwards = [&](Group& gr) -> int {
        gr.fw.resize(2 * gr.nb);
        for (int b = 0; b < gr.nb; ++b) {
            const int k = gr.k0 + b, r0 = sh->rec_off[k], nr = sh->rec_off[k + 1] - r0;
            ShotSpec sp;
            sp.mode = MODE_FWD;
            sp.nt = nt;
            sp.w = w;
            sp.zs = sh->zs[k];
            sp.xs = sh->xs[k];
            sp.nr = nr;
            sp.zr = sh->zr + r0;
            sp.xr = sh->xr + r0;
            CHK(make_shot(ctx, mt, sp, gr.fw[2 * b]));          // observed: true model
            sp.T = T;
            sp.K = K;
            sp.ncol = ncol;
            sp.snap_fmt = snap_fmt;
            sp.ring_order = ctx->ring_order;
            CHK(make_shot(ctx, ms, sp, gr.fw[2 * b + 1]));      // background + snapshots / ring
        }
        return MNAD_OK;
    };
    // after the forwards ran: gathers out, d_obs - d_syn and the adjoint shots
    auto adjoints = [&](Group& gr) -> int {
        gr.ad.resize(gr.nb);
        gr.rc.resize(ring ? gr.nb : 0);
        for (int b = 0; b < gr.nb; ++b) {
            const int k = gr.k0 + b, r0 = sh->rec_off[k], nr = sh->rec_off[k + 1] - r0;
            if (dobs) CHK(download_gather(ctx, gr.fw[2 * b], nt, nr, dobs + (size_t)r0 * nt));
            if (dsyn) CHK(download_gather(ctx, gr.fw[2 * b + 1], nt, nr, dsyn + (size_t)r0 * nt));
            gr.fw[2 * b].release_state();
            if (!ring) gr.fw[2 * b + 1].release_state();   // ring: the reconstruction starts there
            ShotSpec sp;
            sp.mode = MODE_ADJ;
            sp.nt = nt;
            sp.nr = nr;
            sp.zr = sh->zr + r0;
            sp.xr = sh->xr + r0;
            sp.T = T;
            sp.K = K;
            sp.ncol = ncol;
            CHK(make_shot(ctx, ms, sp, gr.ad[b]));
            if (gr.ad[b].nu > 0)
                HIPCHK(mnad_launch_make_radj(gr.fw[2 * b].gather.as<float>(),
                                             gr.fw[2 * b + 1].gather.as<float>(),
                                             gr.ad[b].arec.as<float>(), nt, gr.ad[b].nu,
                                             gr.ad[b].radj.as<float>(), ctx->stream));
            if (ring)
                CHK(make_rec_shot(ctx, gr.fw[2 * b + 1], gr.ad[b], gr.rc[b]));
            else
                attach_snaps(gr.ad[b], std::move(gr.fw[2 * b + 1].snaps), snap_fmt,
                             gr.fw[2 * b + 1].d.sscale);
        }
        return MNAD_OK;
    };
    // after the adjoints ran: add their images and illuminations to the stack
    auto stack = [&](Group& gr) -> int {
        for (int b = 0; b < gr.nb; ++b)
            HIPCHK(mnad_launch_stack(gr.ad[b].img.as<float>(), gr.ad[b].imgc.as<float>(),
                                     gr.fw[2 * b + 1].illum.as<float>(), (float)T, (long)ni,
                                     st.isum.as<float>(), st.isumc.as<float>(),
                                     st.lsum.as<float>(), st.lsumc.as<float>(), ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        st.nshots += gr.nb;
        return MNAD_OK;
    };

    if (!pipelined) {   // in order; serial: one propagation per launch sequence
        for (int k0 = 0; k0 < sh->nshots; k0 += B) {
            Group gr;
            gr.k0 = k0;
            gr.nb = std::min(sh->nshots, k0 + B) - k0;
            CHK(forwards(gr));
            std::vector<Shot*> run;
            for (Shot& s1 : gr.fw) run.push_back(&s1);
            if (serial) {
                for (Shot* s1 : run) CHK(run_batch(ctx, std::vector<Shot*>{s1}, flags));
            } else {
                CHK(run_batch(ctx, run, flags));
            }
            CHK(adjoints(gr));
            run.clear();
            if (serial) {   // an adjoint and its reconstruction share their launches
                for (int b = 0; b < gr.nb; ++b) {
                    std::vector<Shot*> one{&gr.ad[b]};
                    if (ring) one.push_back(&gr.rc[b]);
                    CHK(run_batch(ctx, one, flags));
                }
            } else {
                for (Shot& s1 : gr.ad) run.push_back(&s1);
                for (Shot& s1 : gr.rc) run.push_back(&s1);
                CHK(run_batch(ctx, run, flags));
            }
            CHK(stack(gr));
        }
        return MNAD_OK;
    }
    std::unique_ptr<Group> prev;   // forwards done, adjoints still to run
    for (int k0 = 0; k0 < sh->nshots || prev; k0 += B) {
        std::unique_ptr<Group> cur;
        std::vector<Shot*> run;
        if (k0 < sh->nshots) {
            cur.reset(new Group);
            cur->k0 = k0;
            cur->nb = std::min(sh->nshots, k0 + B) - k0;
            CHK(forwards(*cur));
            for (Shot& s1 : cur->fw) run.push_back(&s1);
        }
        if (prev) {
            CHK(adjoints(*prev));
            for (Shot& s1 : prev->ad) run.push_back(&s1);
            for (Shot& s1 : prev->rc) run.push_back(&s1);
        }
        CHK(run_batch(ctx, run, flags));
        if (prev) CHK(stack(*prev));
        prev = std::move(cur);
    }
    return MNAD_OK;
}

extern "C" int mnad_rtm_schedule(mnad_ctx* ctx, int* shots_per_group, int* pipelined)
{
    ENTER(ctx);
    if (shots_per_group) *shots_per_group = ctx->last_group;
    if (pipelined) *pipelined = ctx->last_pipelined;
    return MNAD_OK;
}

extern "C" int mnad_fwd_shots(mnad_ctx* ctx, const mnad_grid* g, const double* v, int nt,
                              const double* w, const mnad_shots* sh, int batch, int flags,
                              double* datau)
{
    ENTER(ctx);
    CHK(check_shots(sh));
    Model m;
    CHK(make_model(ctx, g, v, m));
    int maxnr = 0;
    for (int k = 0; k < sh->nshots; ++k) maxnr = std::max(maxnr, sh->rec_off[k + 1] - sh->rec_off[k]);
    const TunePick* tp = nullptr;
    if (!(flags & MNAD_FLAG_SERIAL)) {
        tp = find_tuned(ctx, *g);
        if (!tp && batch == 0 && ((flags & MNAD_FLAG_AUTOTUNE) || ctx->autotune_on_use)) {
            TunePick pick;
            CHK(autotune(ctx, m, 32, &pick));
            tp = find_tuned(ctx, *g);
        }
    }
    TuningGuard tg(ctx, tp);
    if (batch == 0 && tp) batch = tp->batch;
    if (flags & MNAD_FLAG_SERIAL) batch = 1;
    const int B = pick_batch(ctx, shot_bytes(m, nt, maxnr), sh->nshots, batch);
    for (int k0 = 0; k0 < sh->nshots; k0 += B) {
        const int k1 = std::min(sh->nshots, k0 + B), nb = k1 - k0;
        std::vector<Shot> fw(nb);
        std::vector<Shot*> run;
        for (int b = 0; b < nb; ++b) {
            const int k = k0 + b, r0 = sh->rec_off[k];
            ShotSpec sp;
            sp.mode = MODE_FWD;
            sp.nt = nt;
            sp.w = w;
            sp.zs = sh->zs[k];
            sp.xs = sh->xs[k];
            sp.nr = sh->rec_off[k + 1] - r0;
            sp.zr = sh->zr + r0;
            sp.xr = sh->xr + r0;
            CHK(make_shot(ctx, m, sp, fw[b]));
            run.push_back(&fw[b]);
        }
        CHK(run_batch(ctx, run, flags));
        for (int b = 0; b < nb; ++b) {
            const int k = k0 + b, r0 = sh->rec_off[k];
            if (datau)
                CHK(download_gather(ctx, fw[b], nt, sh->rec_off[k + 1] - r0, datau + (size_t)r0 * nt));
        }
    }
    return MNAD_OK;
}
