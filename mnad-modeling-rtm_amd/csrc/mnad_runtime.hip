// mnad_runtime.hip -- host runtime behind include/mnad.h.
//
// Owns the device context (stream, caching memory pool, kept source
// wavefields, image stacks), builds per-shot device descriptors, and drives
// the step kernel over batches of shots: one launch per time step covers
// every tile of every shot in the batch (the reference's parfor over shots,
// main_Acoustic_Data.m:169-171, main_Acoustic_RTM.m:223-226,250-254, becomes
// the batch axis), replayed in hipGraph chunks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "mnad.h"
#include "mnad_device.h"
#include "mnad_internal.h"

namespace mnad {

extern "C" {
hipError_t mnad_launch_setup_model(const double* v, int nz, int nx, int pitch, int npml,
                                   double dx, double dz, double dt, float* rx, float* cfx,
                                   float* cfz, hipStream_t st);
hipError_t mnad_launch_step(const ShotDev* shots, const TileRef* tiles, int ntiles,
                            const int* base, int rel, int kernel, int rows, int brows,
                            hipStream_t st);
hipError_t mnad_launch_step2q(const ShotDev* shots, int nshots, const TileRef* tiles, int ntiles,
                              const int* base, int rel, int rows, int brows, int k1, int passes,
                              hipStream_t st);
hipError_t mnad_launch_advance(int* base, int n, hipStream_t st);
hipError_t mnad_launch_init_source(ShotDev* shots, int nshots, hipStream_t st);
hipError_t mnad_launch_make_radj(const float* dobs, const float* dsyn, const float* arec,
                                 int nt, int nrec, float* radj, hipStream_t st);
hipError_t mnad_launch_stack(const float* img, const float* imgc, const float* illum, float T,
                             long n, float* isum, float* isumc, float* lsum, float* lsumc,
                             hipStream_t st);
hipError_t mnad_launch_fill(float* p, long n, float val, hipStream_t st);
int mnad_resident_max_rows();
hipError_t mnad_resident_fit(int bh, int* per_cu);
long mnad_resident_edge_records(int bh);
hipError_t mnad_launch_resident(const ShotDev* shots, const ResBlock* blocks, int nblocks,
                                unsigned long long* ebuf, int* err, int nsteps, int bh, int spin,
                                hipStream_t st);
}

static thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

const char* last_error() { return g_err.c_str(); }

#define HIPCHK(x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess)                                                          \
            return fail(e_ == hipErrorOutOfMemory ? MNAD_ERR_NOMEM : MNAD_ERR_HIP,     \
                        "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                     \
    } while (0)

#define CHK(x)                  \
    do {                        \
        int rc_ = (x);          \
        if (rc_ != MNAD_OK)     \
            return rc_;         \
    } while (0)

// ------------------------------------------------------------------ pool
void* Pool::get(size_t n, int* rc)
{
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(n);
    if (it != cache.end()) {
        void* p = it->second;
        cache.erase(it);
        cached -= n;
        in_use += n;
        *rc = MNAD_OK;
        return p;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        // Free everything cached and retry once.
        for (auto& kv : cache) (void)hipFree(kv.second);
        cache.clear();
        cached = 0;
        e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            *rc = fail(MNAD_ERR_NOMEM, "device allocation of %zu bytes failed (%zu in use)", n,
                       in_use);
            return nullptr;
        }
    }
    in_use += n;
    *rc = MNAD_OK;
    return p;
}

void Pool::put(void* p, size_t n)
{
    std::lock_guard<std::mutex> lk(mu);
    cache.emplace(n, p);
    in_use -= n;
    cached += n;
}

void Pool::trim()
{
    std::lock_guard<std::mutex> lk(mu);
    for (auto& kv : cache) (void)hipFree(kv.second);
    cache.clear();
    cached = 0;
}

void DevBuf::reset()
{
    if (p && pool) pool->put(p, n);
    p = nullptr;
    n = 0;
    pool = nullptr;
}

int dalloc(mnad_ctx* c, DevBuf& b, size_t bytes, bool zero)
{
    b.reset();
    if (bytes == 0) return MNAD_OK;
    bytes = (bytes + 255) & ~size_t(255);
    int rc;
    void* p = c->pool.get(bytes, &rc);
    if (!p) return rc;
    b.p = p;
    b.n = bytes;
    b.pool = &c->pool;
    if (zero) HIPCHK(hipMemsetAsync(p, 0, bytes, c->stream));
    return MNAD_OK;
}

int copy_in(mnad_ctx* c, void* dst, const void* src, size_t bytes)
{
    if (bytes == 0) return MNAD_OK;
    // A slot is free once the copy that last read it has run.  A caller that
    // queues several uploads behind work already on the stream (the serving
    // loop's set-up of new requests while a chunk runs) takes a free slot
    // that is large enough, else the next slot in turn, and waits only when
    // that one's copy has not run yet.
    int pick = -1;
    for (int k = 0; k < mnad_ctx::kStageSlots && pick < 0; ++k) {
        const int q = (c->stage_next + k) % mnad_ctx::kStageSlots;
        mnad_ctx::StageSlot& x = c->stage[q];
        if (x.n < bytes) continue;
        if (x.pending) {
            const hipError_t e = hipEventQuery(x.ev);
            if (e == hipErrorNotReady) continue;
            HIPCHK(e);
            x.pending = false;
        }
        pick = q;
    }
    if (pick < 0) pick = c->stage_next;
    c->stage_next = (pick + 1) % mnad_ctx::kStageSlots;
    mnad_ctx::StageSlot& sl = c->stage[pick];
    if (sl.pending) {
        HIPCHK(hipEventSynchronize(sl.ev));
        sl.pending = false;
    }
    if (bytes > sl.n) {
        if (sl.p) (void)hipHostFree(sl.p);
        sl.p = nullptr;
        sl.n = 0;
        const size_t n = std::max(bytes, (size_t)1 << 20);
        if (hipHostMalloc(&sl.p, n, hipHostMallocDefault) != hipSuccess) {
            // no pinned memory: copy from the caller's buffer and wait for it
            (void)hipGetLastError();
            sl.p = nullptr;
            HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            return MNAD_OK;
        }
        sl.n = n;
    }
    if (!sl.ev) HIPCHK(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
    std::memcpy(sl.p, src, bytes);
    HIPCHK(hipMemcpyAsync(dst, sl.p, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(sl.ev, c->stream));
    sl.pending = true;
    return MNAD_OK;
}

int upload(mnad_ctx* c, DevBuf& b, const void* src, size_t bytes)
{
    CHK(dalloc(c, b, bytes, false));
    return copy_in(c, b.p, src, bytes);   // the caller's buffer is free on return
}

void release_batch_cache(mnad_ctx* c)
{
    for (GraphEntry& e : c->graphs) {
        (void)hipGraphExecDestroy(e.exec);
        (void)hipGraphDestroy(e.graph);
    }
    c->graphs.clear();
    c->bdesc.reset();
    c->btab.reset();
    c->bbase.reset();
    c->bedge.reset();
    for (mnad_ctx::StageSlot& sl : c->stage) {
        if (sl.pending) (void)hipEventSynchronize(sl.ev);
        sl.pending = false;
        if (sl.p) (void)hipHostFree(sl.p);
        sl.p = nullptr;
        sl.n = 0;
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        sl.ev = nullptr;
    }
    c->stage_next = 0;
}

// Grows a persistent batch buffer; graphs that read the old one are dropped.
static int ensure_batch_buf(mnad_ctx* c, DevBuf& b, size_t bytes)
{
    if (b.n >= bytes && b.p) return MNAD_OK;
    const void* old = b.p;
    for (size_t k = 0; k < c->graphs.size();) {
        GraphEntry& e = c->graphs[k];
        if (old && (e.ds == old || e.dt == old || e.base == old)) {
            (void)hipGraphExecDestroy(e.exec);
            (void)hipGraphDestroy(e.graph);
            c->graphs.erase(c->graphs.begin() + k);
        } else {
            ++k;
        }
    }
    return dalloc(c, b, std::max(bytes, 2 * b.n), false);
}

// ----------------------------------------------------------------- model
int check_grid(const mnad_grid* g)
{
    if (!g) return fail(MNAD_ERR_ARG, "grid is NULL");
    if (g->nz < 3 || g->nx < 3 || g->npml < 0)
        return fail(MNAD_ERR_ARG, "grid %dx%d (npml %d) too small", g->nz, g->nx, g->npml);
    if (g->nx < 2 * g->npml + 1 || g->nz < 2 * g->npml + 1)
        return fail(MNAD_ERR_ARG,
                    "grid %dx%d cannot hold CPML strips of %d cells on each side "
                    "(C_cof.m:6,38 needs nx, nz >= 2*d_pml)",
                    g->nz, g->nx, g->npml);
    if (!(g->dx > 0) || !(g->dz > 0) || !(g->dt > 0) || !std::isfinite(g->dx) ||
        !std::isfinite(g->dz) || !std::isfinite(g->dt))
        return fail(MNAD_ERR_ARG, "dx, dz, dt must be positive and finite");
    if ((long)g->nz * (long)g->nx > (1L << 30))
        return fail(MNAD_ERR_ARG, "grid of %ld points exceeds the 2^30 limit",
                    (long)g->nz * g->nx);
    return MNAD_OK;
}

// Validation of a velocity model (n >= 1 values): every value positive and
// finite, and the range.  A positive finite double has bits in
// [1, 0x7FEFFFFFFFFFFFFF], and on that range the order of the bit patterns is
// the order of the values, so the pass runs on integers in independent lanes
// that the compiler vectorises.  Returns n, or the index of the first bad value.
size_t scan_model(const double* v, size_t n, double* vmin, double* vmax)
{
    constexpr int L = 8;
    int64_t lo[L], hi[L];
    uint64_t bad = 0;
    for (int i = 0; i < L; ++i) {
        lo[i] = INT64_MAX;
        hi[i] = 0;
    }
    size_t k = 0;
    for (; k + L <= n; k += L)
        for (int i = 0; i < L; ++i) {
            int64_t b;
            std::memcpy(&b, v + k + i, sizeof b);
            bad |= ((uint64_t)b - 1u) >= 0x7FEFFFFFFFFFFFFFull;
            lo[i] = b < lo[i] ? b : lo[i];
            hi[i] = b > hi[i] ? b : hi[i];
        }
    for (; k < n; ++k) {
        int64_t b;
        std::memcpy(&b, v + k, sizeof b);
        bad |= ((uint64_t)b - 1u) >= 0x7FEFFFFFFFFFFFFFull;
        lo[0] = b < lo[0] ? b : lo[0];
        hi[0] = b > hi[0] ? b : hi[0];
    }
    if (bad)
        for (size_t q = 0; q < n; ++q)
            if (!(v[q] > 0.0 && v[q] < INFINITY)) return q;
    for (int i = 1; i < L; ++i) {
        lo[0] = std::min(lo[0], lo[i]);
        hi[0] = std::max(hi[0], hi[i]);
    }
    std::memcpy(vmin, &lo[0], sizeof *vmin);
    std::memcpy(vmax, &hi[0], sizeof *vmax);
    return n;
}

// 128-bit digest of a model's FP64 bit patterns, the model-cache key.  Two
// multiply-xorshift chains (one fed by xor, one by addition), each in four
// interleaved lanes that take two values a step as w0 + rotl(w1, 29); the
// lanes are folded in order at the end.  Each step is a bijection of the lane
// state for a given input and of either value for the other, so changing any
// single value changes both halves of the digest.
void model_digest(const double* v, size_t n, uint64_t* da, uint64_t* db)
{
    constexpr int L = 4;
    auto rotl = [](uint64_t x, int r) { return (x << r) | (x >> (64 - r)); };
    auto step_a = [](uint64_t h, uint64_t w) {
        h = (h ^ w) * 0x100000001B3ull;
        return h ^ (h >> 29);
    };
    auto step_b = [](uint64_t h, uint64_t w) {
        h = (h + w) * 0x9E3779B97F4A7C15ull;
        return h ^ (h >> 31);
    };
    uint64_t ha[L], hb[L];
    for (int i = 0; i < L; ++i) {
        ha[i] = 0x243F6A8885A308D3ull ^ (0x9E3779B97F4A7C15ull * (uint64_t)(i + 1));
        hb[i] = 0x13198A2E03707344ull + 0xC2B2AE3D27D4EB4Full * (uint64_t)(i + 1);
    }
    size_t k = 0;
    for (; k + 2 * L <= n; k += 2 * L)
        for (int i = 0; i < L; ++i) {
            uint64_t w0, w1;
            std::memcpy(&w0, v + k + i, sizeof w0);
            std::memcpy(&w1, v + k + L + i, sizeof w1);
            const uint64_t w = w0 + rotl(w1, 29);
            ha[i] = step_a(ha[i], w);
            hb[i] = step_b(hb[i], w);
        }
    for (; k < n; ++k) {
        uint64_t w;
        std::memcpy(&w, v + k, sizeof w);
        ha[k % L] = step_a(ha[k % L], w);
        hb[k % L] = step_b(hb[k % L], w);
    }
    for (int i = 1; i < L; ++i) {
        ha[0] = step_a(ha[0], ha[i]);
        hb[0] = step_b(hb[0], hb[i]);
    }
    *da = ha[0];
    *db = hb[0];
}

int make_model(mnad_ctx* c, const mnad_grid* g, const double* v, Model& m)
{
    CHK(check_grid(g));
    if (!v) return fail(MNAD_ERR_ARG, "velocity is NULL");
    const size_t n = (size_t)g->nz * g->nx;
    double vmin, vmax;
    const size_t bad = scan_model(v, n, &vmin, &vmax);
    if (bad < n)
        return fail(MNAD_ERR_ARG, "velocity must be finite and positive (v[%zu] = %g)", bad, v[bad]);
    uint64_t ha, hb;
    model_digest(v, n, &ha, &hb);
    const double courant = vmax * g->dt / std::min(g->dx, g->dz);
    if (courant > 1.0)
        return fail(MNAD_ERR_ARG,
                    "vmax*dt/min(dx,dz) = %.3f > 1: unstable (the reference picks "
                    "dt = 0.8*0.65*dx/vmax, main_Acoustic_Data.m:71,77)",
                    courant);
    m.g = *g;
    m.vmin = vmin;
    m.vmax = vmax;
    m.vhost = v;
    m.pitch = ((MNAD_XO + g->nx + 2) + 31) / 32 * 32;
    m.rows = g->nz + 2 * MNAD_ZO + MNAD_ZPAD;
    m.elems = (size_t)m.rows * m.pitch;
    const ModelKey key{g->nz, g->nx, g->npml, g->dx, g->dz, g->dt, ha, hb};
    for (size_t k = 0; k < c->models.size(); ++k) {
        if (c->models[k].key == key) {
            ModelEntry e = c->models[k];
            c->models.erase(c->models.begin() + k);
            c->models.push_back(e);
            m.bufs = e.bufs;
            return MNAD_OK;
        }
    }
    auto bufs = std::make_shared<ModelBufs>();
    DevBuf dv;
    CHK(upload(c, dv, v, n * sizeof(double)));
    CHK(dalloc(c, bufs->rx, m.elems * sizeof(float), true));
    const int np = g->npml;
    if (np > 0) {
        CHK(dalloc(c, bufs->cfx, 3 * (size_t)g->nz * 2 * np * sizeof(float), true));
        // z strips in rows of `pitch` floats: lane accesses stay 128-byte aligned
        CHK(dalloc(c, bufs->cfz, 3 * (size_t)2 * np * m.pitch * sizeof(float), true));
    }
    HIPCHK(mnad_launch_setup_model(dv.as<double>(), g->nz, g->nx, m.pitch, np, g->dx, g->dz,
                                   g->dt, bufs->rx.as<float>(), bufs->cfx.as<float>(),
                                   bufs->cfz.as<float>(), c->stream));
    // dv goes back to the pool when this returns; the stream order keeps the
    // set-up kernel ahead of any later reuse of that memory.
    m.bufs = bufs;
    c->models.push_back(ModelEntry{key, bufs});
    if (c->models.size() > kModelCache) c->models.erase(c->models.begin());
    return MNAD_OK;
}

// ------------------------------------------------------------------ shot
static size_t snap_elem(int fmt) { return fmt == MNAD_SNAP_F32 ? 4 : 2; }

// Unique receiver points of a shot (s.uid, s.upts, s.nu) and the row/column
// lookup tables of the kernel (rrow [nz], rcol [rows with receivers * nx]);
// host only.
int receiver_tables(const mnad_grid& g, int nr, const int* zr, const int* xr, Shot& s,
                    std::vector<int>& rrow, std::vector<int>& rcol)
{
    s.uid.assign(nr, -1);
    s.upts.clear();
    std::map<std::pair<int, int>, int> seen;
    for (int r = 0; r < nr; ++r) {
        if (zr[r] < 0 || zr[r] >= g.nz || xr[r] < 0 || xr[r] >= g.nx)
            return fail(MNAD_ERR_ARG, "receiver %d at (%d,%d) outside the %dx%d grid", r, zr[r],
                        xr[r], g.nz, g.nx);
        auto key = std::make_pair(zr[r], xr[r]);
        auto it = seen.find(key);
        if (it == seen.end()) {
            int u = (int)s.upts.size();
            seen.emplace(key, u);
            s.upts.push_back(key);
            s.uid[r] = u;
        } else {
            s.uid[r] = it->second;
        }
    }
    s.nu = (int)s.upts.size();
    s.d.nrec = s.nu;
    rrow.assign(g.nz, -1);
    rcol.clear();
    s.d.rlo = 1;
    s.d.rhi = 0;
    if (s.nu == 0) return MNAD_OK;
    std::vector<int> rows;
    for (auto& pt : s.upts)
        if (rrow[pt.first] < 0) {
            rrow[pt.first] = (int)rows.size() * g.nx;
            rows.push_back(pt.first);
        }
    s.d.rlo = *std::min_element(rows.begin(), rows.end());
    s.d.rhi = *std::max_element(rows.begin(), rows.end());
    rcol.assign(rows.size() * (size_t)g.nx, -1);
    for (int u = 0; u < s.nu; ++u) rcol[rrow[s.upts[u].first] + s.upts[u].second] = u;
    return MNAD_OK;
}

// receiver_tables on the device, and a = (v dt)^2 at each unique receiver.
static int build_receivers(mnad_ctx* c, const Model& m, int nr, const int* zr, const int* xr,
                           Shot& s)
{
    const mnad_grid& g = m.g;
    std::vector<int> rrow, rcol;
    CHK(receiver_tables(g, nr, zr, xr, s, rrow, rcol));
    if (s.nu == 0) return MNAD_OK;
    CHK(upload(c, s.rrow, rrow.data(), rrow.size() * sizeof(int)));
    CHK(upload(c, s.rcol, rcol.data(), rcol.size() * sizeof(int)));
    s.d.rrow = s.rrow.as<int>();
    s.d.rcol = s.rcol.as<int>();
    // a = (v dt)^2 at each unique receiver point (Born_adj_MNAD.m:36,110).
    s.arec_h.resize(s.nu);
    for (int u = 0; u < s.nu; ++u) {
        const double vv = m.vhost[(size_t)s.upts[u].first + (size_t)s.upts[u].second * g.nz];
        s.arec_h[u] = (vv * g.dt) * (vv * g.dt);
    }
    return MNAD_OK;
}

void base_desc(const mnad_grid& g, int pitch, int nt, int mode, int nlev, ShotDev& d)
{
    d = ShotDev{};
    d.nz = g.nz;
    d.nx = g.nx;
    d.pitch = pitch;
    d.npml = g.npml;
    d.nt = nt;
    d.mode = mode;
    d.kap = (float)((g.dx * g.dx) / (g.dz * g.dz));
    incr_consts((g.dx * g.dx) / (g.dz * g.dz), d.ic);
    d.kap1 = (g.dx * g.dx) / (g.dz * g.dz) == 1.0;
    d.inv_dt2 = (float)(1.0 / (g.dt * g.dt));
    d.ntx = (g.nx + 63) / 64;
    d.src_j = d.src_i = -1;
    d.nlev = nlev;
    d.rlo = 1;   // no receivers
    d.rhi = 0;
    d.pbytes = 0;
}

int make_shot(mnad_ctx* c, const Model& m, const ShotSpec& sp, Shot& s)
{
    const mnad_grid& g = m.g;
    if (sp.nt < 1) return fail(MNAD_ERR_ARG, "nt must be >= 1");
    base_desc(g, m.pitch, sp.nt, sp.mode, two_step_kernel(c->kernel) ? 4 : 2, s.d);
    s.d.rx = m.rx();
    // all 3 nlev field planes in one allocation, plane k = 3 level + field at
    // k * pbytes (step kernel 13 addresses them through one buffer resource)
    const size_t pbytes = plane_stride_bytes(g.nz, m.pitch);
    CHK(dalloc(c, s.state[0], 3 * s.d.nlev * pbytes, true));
    s.d.pbytes = (unsigned)pbytes;
    for (int k = 0; k < 3 * s.d.nlev; ++k)
        s.d.f[k / 3][k % 3] = s.state[0].as<float>() + k * (pbytes / sizeof(float));
    const int np = g.npml;
    if (np > 0) {
        // two-steps-per-launch kernels keep two sets of memories (read one, write the other)
        const size_t planes = s.d.nlev == 4 ? 18 : 9;
        CHK(dalloc(c, s.psx, planes * g.nz * 2 * np * sizeof(float), true));
        CHK(dalloc(c, s.psz, planes * 2 * np * m.pitch * sizeof(float), true));
        s.d.psx = s.psx.as<float>();
        s.d.psz = s.psz.as<float>();
        s.d.cfx = m.cfx();
        s.d.cfz = m.cfz();
    }
    CHK(build_receivers(c, m, sp.nr, sp.zr, sp.xr, s));
    const size_t ni = (size_t)(g.nz - 2 * np) * (g.nx - 2 * np);
    if (sp.mode == MODE_FWD) {
        if (sp.zs < 0 || sp.zs >= g.nz || sp.xs < 0 || sp.xs >= g.nx)
            return fail(MNAD_ERR_ARG, "source (%d,%d) outside the %dx%d grid", sp.zs, sp.xs,
                        g.nz, g.nx);
        if (!sp.w) return fail(MNAD_ERR_ARG, "wavelet is NULL");
        s.d.src_j = sp.zs;
        s.d.src_i = sp.xs;
        // a_src * w(tt): fwd_PDE_exact_MNAD.m:39,111 (FP64 on the host).
        const double vs = m.vhost[(size_t)sp.zs + (size_t)sp.xs * g.nz];
        const double asrc = vs * vs * (g.dt * g.dt);
        std::vector<float> amp(sp.nt + 1, 0.0f);
        double wmax = 0;
        for (int t = 0; t < sp.nt; ++t) {
            if (!std::isfinite(sp.w[t])) return fail(MNAD_ERR_ARG, "wavelet sample %d is not finite", t);
            amp[t] = (float)(asrc * sp.w[t]);
            wmax = std::max(wmax, std::fabs(sp.w[t]));
        }
        CHK(upload(c, s.amp, amp.data(), amp.size() * sizeof(float)));
        s.d.src_amp = s.amp.as<float>();
        if (s.nu > 0) {
            CHK(dalloc(c, s.gather, (size_t)sp.nt * s.nu * sizeof(float), true));
            s.d.gather = s.gather.as<float>();
        }
        if (sp.T > 0 && ring_format(sp.snap_fmt)) {
            // illumination every T steps, boundary ring every step, no snapshots
            if (np < 2) return fail(MNAD_ERR_ARG, "ring storage needs npml >= 2 (npml = %d)", np);
            if (sp.ring_order < 2 || sp.ring_order > RING_MAX_ORDER)
                return fail(MNAD_ERR_ARG, "ring predictor order %d not in 2..%d", sp.ring_order,
                            RING_MAX_ORDER);
            s.d.T = sp.T;
            s.d.K = sp.K;
            s.d.shift = sp.ncol - sp.K;
            s.d.sfmt = SNAP_F32;
            s.d.snaps = nullptr;
            CHK(dalloc(c, s.illum, ni * sizeof(float), true));
            s.d.illum = s.illum.as<float>();
            const int P = sp.ring_order, nr = ring_cells(g.nz, g.nx, np);
            s.d.rord = P;
            s.d.nring = nr;
            s.d.rfmt = sp.snap_fmt == MNAD_SNAP_RING_F16 ? SNAP_F16 : SNAP_BF16;
            // fixed-point unit 2^-E = 2^-40 of the source amplitude a_src max|w|;
            // the field stays far below 2^22 times that (|G^| < 2^62)
            const double amax = asrc * std::max(wmax, 1e-300);
            const int E = 40 - (int)std::ceil(std::log2(amax));
            s.d.rsc = std::ldexp(1.0, E);
            s.d.rsi = std::ldexp(1.0, -E);
            // F16 codes: a residual of 2^P times the source amplitude maps to 2^14
            s.d.rcs = s.d.rfmt == SNAP_F16 ? std::ldexp(1.0, -(26 + P)) : 1.0;
            s.d.rci = 1.0 / s.d.rcs;
            CHK(dalloc(c, s.rq, (size_t)sp.nt * 3 * nr * sizeof(uint16_t), false));
            CHK(dalloc(c, s.rst, (size_t)P * 3 * nr * sizeof(int64_t), true));
            s.d.rq = s.rq.as<unsigned short>();
            s.d.rst = s.rst.as<long long>();
        } else if (sp.T > 0) {
            s.d.T = sp.T;
            s.d.K = sp.K;
            s.d.shift = sp.ncol - sp.K;
            s.d.sfmt = sp.snap_fmt;
            // FP16: power-of-two scale mapping the largest expected d2u (the
            // source point, ~4 a_src max|w| / dt^2) to 2^12, far below 65504.
            double est = 4.0 * asrc * std::max(wmax, 1e-30) / (g.dt * g.dt);
            double sc = std::ldexp(1.0, (int)std::floor(std::log2(4096.0 / est)));
            s.d.sscale = (float)sc;
            s.d.sinv = (float)(1.0 / sc);
            CHK(dalloc(c, s.snaps, (size_t)sp.K * ni * snap_elem(sp.snap_fmt), false));
            CHK(dalloc(c, s.illum, ni * sizeof(float), true));
            s.d.snaps = s.snaps.p;
            s.d.illum = s.illum.as<float>();
        }
    } else {
        if (s.nu > 0) {
            CHK(dalloc(c, s.radj, (size_t)sp.nt * s.nu * sizeof(float), true));
            s.d.radj = s.radj.as<float>();
            std::vector<float> ah(s.arec_h.begin(), s.arec_h.end());
            CHK(upload(c, s.arec, ah.data(), ah.size() * sizeof(float)));
        }
        if (sp.T > 0) {
            s.d.T = sp.T;
            s.d.K = sp.K;
            s.d.shift = sp.ncol - sp.K;
            CHK(dalloc(c, s.img, ni * sizeof(float), true));
            CHK(dalloc(c, s.imgc, ni * sizeof(float), true));
            s.d.img = s.img.as<float>();
            s.d.imgc = s.imgc.as<float>();
        }
    }
    return MNAD_OK;
}

int make_rec_shot(mnad_ctx* c, Shot& fwd, Shot& adj, Shot& rec)
{
    const ShotDev& f = fwd.d;
    if (f.mode != MODE_FWD || f.rord <= 0 || f.T <= 0)
        return fail(MNAD_ERR_STATE, "reverse reconstruction needs a forward that saved its ring");
    if (f.nlev != 2 && f.nlev != 4)
        return fail(MNAD_ERR_STATE, "reverse reconstruction needs a forward of 2 or 4 levels");
    rec.d = f;
    ShotDev& d = rec.d;
    d.mode = MODE_REC;
    d.ntf = f.nt;
    // the adjoint images at tt = m T with snapshot column shift + m <= K, m >= 1:
    // the earliest forward step needed is (1 + shift) T
    const int first = (1 + f.shift) * f.T;
    d.nt = std::max(1, f.nt - first + 1);
    d.nrec = 0;
    d.rlo = 1;
    d.rhi = 0;
    d.rrow = nullptr;
    d.rcol = nullptr;
    d.gather = nullptr;
    d.psx = d.psz = nullptr;
    d.illum = nullptr;
    d.img = d.imgc = nullptr;
    d.sfmt = SNAP_SLOT;
    for (int k = 0; k < 6; ++k) rec.state[k] = std::move(fwd.state[k]);
    fwd.psx.reset();
    fwd.psz.reset();
    rec.amp = std::move(fwd.amp);
    rec.rq = std::move(fwd.rq);
    rec.rst = std::move(fwd.rst);
    fwd.d.rq = nullptr;
    fwd.d.rst = nullptr;
    const size_t ni = (size_t)(f.nz - 2 * f.npml) * (f.nx - 2 * f.npml);
    CHK(dalloc(c, rec.snaps, ni * sizeof(float), true));
    d.snaps = rec.snaps.p;
    // the adjoint reads the slot one launch after REC writes it: one-step
    // kernels t0(rec) - t0(adj) = shift T - 1.  Kernel 13 (two steps per
    // launch, the slot written and read by the passes after a launch, the
    // reads first) needs the read L = 2 or 3 steps after the write and both
    // launch origins even (a launch covers steps 2n, 2n + 1 of every shot):
    // shift T - L with L making it even, and L <= T so that the next write
    // comes no earlier than the read's launch
    adj.d.snaps = rec.snaps.p;
    adj.d.sfmt = SNAP_SLOT;
    adj.d.sscale = adj.d.sinv = 1.0f;
    long long lag = (long long)f.shift * f.T - 1;
    if (f.nlev == 4) {
        const int L = (((long long)f.shift * f.T) % 2 == 0) ? 2 : 3;
        if (L > f.T)
            return fail(MNAD_ERR_STATE, "reverse reconstruction on a two-step kernel needs a "
                        "snapshot period of at least 3 (this schedule has %d)", f.T);
        lag = (long long)f.shift * f.T - L;
    }
    if (lag >= 0) {
        adj.t0 = 0;
        rec.t0 = lag;
    } else {
        adj.t0 = -lag;
        rec.t0 = 0;
    }
    return MNAD_OK;
}

// Attach snapshots (from a forward shot or a kept handle) to an adjoint shot.
void attach_snaps(Shot& adj, DevBuf&& snaps, int fmt, float sscale)
{
    adj.snaps = std::move(snaps);
    adj.d.snaps = adj.snaps.p;
    adj.d.sfmt = fmt;
    adj.d.sscale = sscale;
    adj.d.sinv = 1.0f / sscale;
}

// Host-side radj for an adjoint from a [nt x nr] column-major data matrix:
// radj[s][u] = a_u * sum_{r -> u} d(nt - s + 1, r) (1-based time), s >= 1.
int fill_radj_host(mnad_ctx* c, Shot& s, int nt, int nr, const double* data)
{
    if (s.nu == 0) return MNAD_OK;
    std::vector<double> acc((size_t)nt * s.nu, 0.0);
    for (int r = 0; r < nr; ++r) {
        const int u = s.uid[r];
        for (int t = 0; t < nt; ++t) acc[(size_t)t * s.nu + u] += data[(size_t)t + (size_t)r * nt];
    }
    std::vector<float> radj((size_t)nt * s.nu, 0.0f);
    for (int st = 1; st < nt; ++st)
        for (int u = 0; u < s.nu; ++u)
            radj[(size_t)st * s.nu + u] = (float)(s.arec_h[u] * acc[(size_t)(nt - st) * s.nu + u]);
    for (double x : acc)
        if (!std::isfinite(x)) return fail(MNAD_ERR_ARG, "data contains NaN/Inf");
    HIPCHK(hipMemcpyAsync(s.radj.p, radj.data(), radj.size() * sizeof(float),
                          hipMemcpyHostToDevice, c->stream));
    // keep the host buffer alive until the copy is done
    HIPCHK(hipStreamSynchronize(c->stream));
    return MNAD_OK;
}

// ------------------------------------------------------------ batch runner
void tile_shape(const mnad_ctx* c, int* tw, int* th, int* spl)
{
    if (c->kernel == 7) {   // two steps per launch: 60 x (4 rows_per_thread) outputs
        *tw = 60;
        *th = 4 * c->rz;
        *spl = 2;
    } else if (c->kernel == 13) {  // two steps per launch, register queues: block_rows strips of 56
        *tw = 56 * c->brows;       // columns x segments of 8 x rows rows
        *th = 8 * c->rz;
        *spl = 2;
    } else {                // 64 columns x (block_rows * rows_per_thread) rows
                            // (kernel 6: the one-point tiles; see build_tiles)
        *tw = 64;
        *th = c->brows * c->rz;
        *spl = 1;
    }
}

// Interior tiles of step kernels 6 and 10: nc = 4 or 2 columns per lane,
// 62 nc columns and th rows per tile.
static void x4_layout(ShotDev& d, int nc, int nt1)
{
    const int np = d.npml;
    d.nt1 = nt1;
    d.x4lo = np & ~(nc - 1);   // lane 0 of the first tile starts at x4lo - nc: aligned loads
    d.x4hi = d.nx - np;
    const int width = nc * 62;
    if (d.x4hi <= np || d.nz - 2 * np <= 0) {
        d.nx4 = 0;
        d.x4hi = np;            // no interior: the one-point tiles cover everything
    } else {
        d.nx4 = (d.x4hi - d.x4lo + width - 1) / width;
    }
}

int interior_columns_per_lane(int kernel) { return kernel == 6 ? 4 : (kernel == 10 ? 2 : 0); }

void build_tiles(const mnad_ctx* c, std::vector<ShotDev>& h, std::vector<TileRef>& table)
{
    std::vector<TileRef> list;
    std::vector<char> strip;
    std::vector<double> work;   // estimated work of each listed tile, in one-point tiles
    // Two-step tiles are cut by count: weighting their strip tiles measured
    // slower at two shots per launch and made no difference at four or more.
    const double sw = two_step_kernel(c->kernel) ? 1.0 : c->strip_weight;
    int tw, th, spl;
    tile_shape(c, &tw, &th, &spl);
    const int nc = interior_columns_per_lane(c->kernel);
    const bool x4 = nc > 0;
    for (int k = 0; k < (int)h.size(); ++k) {
        const int ntx = (h[k].nx + tw - 1) / tw, ntz = (h[k].nz + th - 1) / th;
        const int np = h[k].npml;
        h[k].ntx = ntx;
        x4_layout(h[k], x4 ? nc : 4, ntx * ntz);
        const int nz = h[k].nz, nx = h[k].nx, x4hi = x4 ? h[k].x4hi : np;
        const bool rec = h[k].mode == MODE_REC;
        for (int t = 0; t < ntx * ntz; ++t) {
            const int tj = t / ntx, ti = t - tj * ntx;
            const int r0 = tj * th, r1 = std::min(nz, r0 + th), c0 = ti * tw,
                      c1 = std::min(nx, c0 + tw);
            // kernels 6, 10: one-point tiles lying wholly in the interior are not listed
            if (x4 && r0 >= np && r1 <= nz - np && c0 >= np && c1 <= x4hi) continue;
            // a reverse reconstruction touches the interior and its two-cell ring only
            if (rec && (r1 <= np - 2 || r0 >= nz - np + 2 || c1 <= np - 2 || c0 >= nx - np + 2))
                continue;
            if (c->kernel == 13) {
                // kernel 13: the tile's work is its interior waves (56 columns
                // passing q_interior_tile) plus its edge waves at their weight
                int nin = 0, ned = 0;
                for (int x0 = c0; x0 < c1; x0 += 56)
                    (q_interior_tile(nx, nz, np, x0, r0, 56, th) ? nin : ned)++;
                list.push_back(TileRef{k, t});
                strip.push_back(ned > 0);
                work.push_back(nin + kQEdgeWeight * ned);
                continue;
            }
            const bool st = !rec && (c0 < np || c1 > nx - np || r0 < np || r1 > nz - np);
            list.push_back(TileRef{k, t});
            strip.push_back(st);
            work.push_back(st ? sw : 1.0);
        }
        if (x4) {
            // a four-column tile: 248 columns at about 2/3 of the one-point cost per
            // point; a two-column tile is taken as half of one
            const int nb = (nz - 2 * np + th - 1) / th;
            for (int u = 0; u < nb * h[k].nx4; ++u) {
                list.push_back(TileRef{k, ntx * ntz + u});
                strip.push_back(0);
                work.push_back(c->x4_weight * nc / 4.0);
            }
        }
    }
    // XCD-aware order: workgroup b is dispatched to XCD b % 8.  The spatial
    // tile list is cut into 8 contiguous runs of equal estimated work and XCD
    // x takes run x, so vertically adjacent tiles share halo rows in the same
    // L2.  Cutting by work rather than by tile count matters because a CPML
    // strip tile costs about twice an interior one (three times the bytes):
    // equal counts hand the top and bottom strips of a shot to two XCDs,
    // whose longer runs then set the launch time (a one-shot launch of the
    // 561 x 1061 grid: 8.7 us by count, 7.8 us by work).  Within each run
    // the strip tiles go first.
    const int nx8 = 8;
    auto order = [&](const std::vector<size_t>& sel) {
        double total = 0.0;
        for (size_t q : sel) total += work[q];
        std::vector<std::vector<size_t>> runs(nx8);
        double acc = 0.0;
        for (size_t q : sel) {
            const double mid = acc + 0.5 * work[q];
            acc += work[q];
            runs[std::min(nx8 - 1, (int)(mid * nx8 / total))].push_back(q);
        }
        size_t per = 0;
        for (const auto& run : runs) per = std::max(per, run.size());
        const size_t at = table.size();
        table.resize(at + per * nx8, TileRef{-1, 0});
        for (int x = 0; x < nx8; ++x) {
            auto& run = runs[x];
            if (c->strip_first)
                std::stable_partition(run.begin(), run.end(), [&](size_t q) { return strip[q] != 0; });
            for (size_t m = 0; m < run.size(); ++m) table[at + m * nx8 + x] = list[run[m]];
        }
    };
    table.clear();
    std::vector<size_t> all(list.size());
    for (size_t q = 0; q < list.size(); ++q) all[q] = q;
    if (!all.empty()) order(all);
}

// Unique HBM bytes a shot's nt steps must move:
// - per padded point: U, Vs, Ws at both levels and rx read, the new level written (40 B);
// - per CPML strip point and direction: nine memories (three fields x three levels) read
//   and written (72 B) and the three damping coefficients c1, h*c2, c3 read (12 B);
// - per receiver: its gather or injection value (4 B);
// - every T steps: the snapshot plus illumination (forward) or image and compensation
//   (adjoint) read and written.
// - saved boundary ring, per ring cell, field and step: the predictor state read
//   and written (8 (P + 1) B) and the 16-bit code written (forward) or read
//   (reverse reconstruction);
// - reverse reconstruction: 40 B per interior point (the CPML strips are not
//   touched), plus the d2u slot written every T steps.
double algorithmic_bytes(const ShotDev& d)
{
    const double N = (double)d.nz * d.nx, np2 = 2.0 * d.npml;
    const double strips = np2 * d.nz + np2 * d.nx;
    const double ni = (double)(d.nz - 2 * d.npml) * (d.nx - 2 * d.npml);
    const double esz = (d.sfmt == SNAP_F32 || d.sfmt == SNAP_SLOT) ? 4.0 : 2.0;
    const double ring = d.rord > 0 ? 3.0 * d.nring * (8.0 * (d.rord + 1) + 2.0) : 0.0;
    if (d.mode == MODE_REC) {
        double b = (double)d.nt * (40.0 * ni + ring + 12.0 * d.nring);   // + decoded ring written
        if (d.T > 0) b += std::floor((double)d.nt / d.T) * ni * 4.0;
        return b;
    }
    double b = (double)d.nt * (40.0 * N + 84.0 * strips + 4.0 * d.nrec + ring);
    if (d.T > 0)
        b += (double)d.K * ni * ((d.snaps || d.mode == MODE_ADJ ? esz : 0.0) +
                                 (d.mode == MODE_FWD ? 8.0 : 16.0));
    return b;
}

// Block height of the resident kernel for a batch: the smallest (at least 4)
// at which the batch takes at most `per_cu` blocks per CU; 0 if none up to
// the kernel's limit does.  *nblocks gets the count at that height.
int resident_rows(const std::vector<ShotDev>& h, int ncu, int per_cu, long* nblocks)
{
    const long cap = (long)std::max(per_cu, 1) * ncu;
    for (int bh = 4; bh <= mnad_resident_max_rows(); ++bh) {
        long n = 0;
        for (const ShotDev& d : h) n += (long)((d.nx + 63) / 64) * ((d.nz + bh - 1) / bh);
        if (n <= cap) {
            *nblocks = n;
            return bh;
        }
    }
    return 0;
}

// Kernel 9: every step of the batch in one cooperative launch of resident
// blocks (64 columns x bh rows of one shot each; c->rz = blocks per CU).
static int run_resident(mnad_ctx* c, std::vector<ShotDev>& h, int nsteps, double bytes)
{
    int ncu = 0;
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
    // the most blocks per CU up to c->rz at which the batch fits on the device
    long total = 0;
    int bh = 0;
    for (int want = std::max(c->rz, 1); want >= 1 && bh == 0; --want) {
        const int b = resident_rows(h, ncu, want, &total);
        if (b == 0) continue;
        int fit = 0;
        HIPCHK(mnad_resident_fit(b, &fit));
        if ((long)fit * ncu >= total) bh = b;
    }
    // MNAD_RESIDENT_ROWS forces a (larger) block height, e.g. to test long
    // blocks on small grids
    if (const char* e = getenv("MNAD_RESIDENT_ROWS")) {
        const int f = atoi(e);
        if (bh > 0 && f > bh && f <= mnad_resident_max_rows()) {
            long n = 0;
            for (const ShotDev& d : h) n += (long)((d.nx + 63) / 64) * ((d.nz + f - 1) / f);
            int fit = 0;
            HIPCHK(mnad_resident_fit(f, &fit));
            if ((long)fit * ncu >= n) {
                bh = f;
                total = n;
            }
        }
    }
    if (bh == 0)
        return fail(MNAD_ERR_ARG, "resident step kernel: the batch does not fit on the device "
                    "as blocks of 64 x %d points or fewer", mnad_resident_max_rows());
    std::vector<ResBlock> blocks;
    blocks.reserve(total);
    for (int k = 0; k < (int)h.size(); ++k) {
        const int nbx = (h[k].nx + 63) / 64, nbz = (h[k].nz + bh - 1) / bh;
        const int first = (int)blocks.size();
        for (int bz = 0; bz < nbz; ++bz)
            for (int bx = 0; bx < nbx; ++bx) {
                ResBlock rb{};
                rb.shot = k;
                rb.r0 = bz * bh;
                rb.c0 = bx * 64;
                int m = 0;
                for (int oz = -1; oz <= 1; ++oz)
                    for (int ox = -1; ox <= 1; ++ox) {
                        if (oz == 0 && ox == 0) continue;
                        const int z = bz + oz, x = bx + ox;
                        rb.nbr[m++] = (z >= 0 && z < nbz && x >= 0 && x < nbx)
                                          ? first + z * nbx + x : -1;
                    }
                blocks.push_back(rb);
            }
    }
    CHK(ensure_batch_buf(c, c->bdesc, h.size() * sizeof(ShotDev)));
    CHK(ensure_batch_buf(c, c->btab, blocks.size() * sizeof(ResBlock)));
    // edge records of every block, then the give-up word
    const long nrec = (long)blocks.size() * mnad_resident_edge_records(bh);
    CHK(ensure_batch_buf(c, c->bedge, (size_t)nrec * 8 + 256));
    CHK(copy_in(c, c->bdesc.p, h.data(), h.size() * sizeof(ShotDev)));
    CHK(copy_in(c, c->btab.p, blocks.data(), blocks.size() * sizeof(ResBlock)));
    auto* ebuf = c->bedge.as<unsigned long long>();
    int* err = reinterpret_cast<int*>(ebuf + nrec);
    // records (level 0: no step waits for it) and the give-up word start at
    // zero (0.0f has all bits clear)
    HIPCHK(mnad_launch_fill(c->bedge.as<float>(), nrec * 2 + 1, 0.0f, c->stream));
    HIPCHK(mnad_launch_init_source(c->bdesc.as<ShotDev>(), (int)h.size(), c->stream));
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    HIPCHK(hipEventCreate(&ev0));
    HIPCHK(hipEventCreate(&ev1));
    struct EvGuard {
        hipEvent_t a, b;
        ~EvGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
    } evg{ev0, ev1};
    HIPCHK(hipEventRecord(ev0, c->stream));
    // polls of one neighbour record before a block gives up (about a second);
    // MNAD_RESIDENT_SPIN lowers it, e.g. to 0 to exercise the give-up path
    int spin = 1 << 20;
    if (const char* e = getenv("MNAD_RESIDENT_SPIN")) spin = std::max(0, atoi(e));
    HIPCHK(mnad_launch_resident(c->bdesc.as<ShotDev>(), c->btab.as<ResBlock>(),
                                (int)blocks.size(), ebuf, err, nsteps, bh, spin, c->stream));
    HIPCHK(hipEventRecord(ev1, c->stream));
    int gave_up = 0;
    HIPCHK(hipMemcpyAsync(&gave_up, err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (gave_up)
        return fail(MNAD_ERR_HIP, "resident step kernel: a block gave up waiting for its "
                    "neighbours; the batch's fields are not valid");
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
    c->t_ms += ms;
    c->t_launches += 1;
    c->t_bytes += bytes;
    return MNAD_OK;
}

int run_batch(mnad_ctx* c, const std::vector<Shot*>& shots, int flags)
{
    const int n = (int)shots.size();
    if (n == 0) return MNAD_OK;
    std::vector<ShotDev> h(n);
    int nsteps = 0;
    const int kern = c->kernel, rz = c->rz, br = c->brows;
    for (int k = 0; k < n; ++k) {
        h[k] = shots[k]->d;
        if (shots[k]->t0 < 0 || shots[k]->t0 > (1 << 20))
            return fail(MNAD_ERR_STATE, "shot launch origin %lld out of range", shots[k]->t0);
        h[k].t0 = (int)shots[k]->t0;
        if (two_step_kernel(kern) && (h[k].t0 & 1))   // its step 0 would fall in a skipped half
            return fail(MNAD_ERR_STATE, "shot launch origin %d is odd: step kernel %d runs two "
                        "steps per launch from even origins", h[k].t0, kern);
        nsteps = std::max(nsteps, h[k].t0 + h[k].nt);
        if (h[k].nlev != (two_step_kernel(kern) ? 4 : 2))
            return fail(MNAD_ERR_STATE, "shot built for %d time levels cannot run on step kernel %d",
                        h[k].nlev, kern);
        if (kern == 13) {   // one buffer resource spans a shot's field planes
            const size_t pe = h[k].pbytes / sizeof(float);
            bool contiguous = h[k].pbytes == plane_stride_bytes(h[k].nz, h[k].pitch) &&
                              3 * h[k].nlev * (size_t)h[k].pbytes < ((size_t)1 << 31);
            for (int q = 0; q < 3 * h[k].nlev; ++q)
                contiguous = contiguous && h[k].f[q / 3][q % 3] == h[k].f[0][0] + q * pe;
            if (!contiguous)
                return fail(MNAD_ERR_STATE, "step kernel 13 needs a shot's field planes in one "
                            "allocation of under 2 GiB; use another kernel for this grid");
        }
        if ((h[k].rord > 0 || h[k].mode == MODE_REC) && !ring_kernel(kern))
            return fail(MNAD_ERR_STATE, "step kernel %d does not save or replay the boundary ring "
                        "(kernels 4, 5, 6, 10, 11, 13 do)", kern);
    }
    if (resident_kernel(kern)) {
        double bytes = 0.0;
        for (int k = 0; k < n; ++k) bytes += algorithmic_bytes(h[k]);
        return run_resident(c, h, nsteps, bytes);
    }
    std::vector<TileRef> table;
    build_tiles(c, h, table);
    // descriptors in the context's persistent buffers, so that a cached graph
    // chunk of the same shape replays on them
    CHK(ensure_batch_buf(c, c->bdesc, h.size() * sizeof(ShotDev)));
    CHK(ensure_batch_buf(c, c->btab, table.size() * sizeof(TileRef)));
    CHK(ensure_batch_buf(c, c->bbase, sizeof(int)));
    CHK(copy_in(c, c->bdesc.p, h.data(), h.size() * sizeof(ShotDev)));
    CHK(copy_in(c, c->btab.p, table.data(), table.size() * sizeof(TileRef)));
    HIPCHK(hipMemsetAsync(c->bbase.p, 0, sizeof(int), c->stream));   // step counter (not captured)
    HIPCHK(mnad_launch_init_source(c->bdesc.as<ShotDev>(), n, c->stream));
    // Algorithmic HBM bytes of the whole sequence (see mnad_timing).
    double bytes = 0.0;
    for (int k = 0; k < n; ++k) bytes += algorithmic_bytes(h[k]);
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    HIPCHK(hipEventCreate(&ev0));
    HIPCHK(hipEventCreate(&ev1));
    struct EvGuard {
        hipEvent_t a, b;
        ~EvGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
    } evg{ev0, ev1};
    const int ntiles = (int)table.size();
    const ShotDev* ds = c->bdesc.as<ShotDev>();
    const TileRef* dt = c->btab.as<TileRef>();
    int* base = c->bbase.as<int>();
    int chunk = c->graph_chunk;
    int tw, th, spl;
    tile_shape(c, &tw, &th, &spl);
    const int nlaunch = (nsteps + spl - 1) / spl;   // a launch covers spl steps
    int k1 = 1;       // kernel 13: every shot has dx == dz
    int passes = 0;   // kernel 13: the work of the passes around each launch
    for (const ShotDev& x : h) {
        k1 &= x.kap1 ? 1 : 0;
        if (x.mode == MODE_REC && x.rord > 0) passes |= 1;
        if ((x.T > 0 && x.mode != MODE_REC) || (x.mode == MODE_FWD && (x.nrec > 0 || x.rord > 0)))
            passes |= 2;
        if (x.T > 0 && x.mode == MODE_REC) passes |= 4;
    }
    auto step = [&](int rel) {
        return kern == 13 ? mnad_launch_step2q(ds, n, dt, ntiles, base, rel, rz, br, k1, passes,
                                               c->stream)
                          : mnad_launch_step(ds, dt, ntiles, base, rel, kern, rz, br, c->stream);
    };
    if ((flags & MNAD_FLAG_NO_GRAPH) || chunk <= 1 || nlaunch < 2 * chunk) {
        HIPCHK(hipEventRecord(ev0, c->stream));
        for (int s = 0; s < nlaunch; ++s) HIPCHK(step(s * spl));
    } else {
        hipGraphExec_t exec = nullptr;
        for (size_t k = 0; k < c->graphs.size(); ++k) {
            const GraphEntry& e = c->graphs[k];
            if (e.kern == kern && e.rz == rz && e.brows == br && e.ntiles == ntiles &&
                e.chunk == chunk && e.spl == spl && e.nshots == n && e.k1 == k1 &&
                e.passes == passes && e.ds == ds && e.dt == dt && e.base == base) {
                exec = e.exec;
                break;
            }
        }
        if (!exec) {
            hipGraph_t graph = nullptr;
            HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            hipError_t ce = hipSuccess;
            for (int r = 0; r < chunk && ce == hipSuccess; ++r)
                ce = step(r * spl);
            if (ce == hipSuccess) ce = mnad_launch_advance(base, chunk * spl, c->stream);
            hipError_t ee = hipStreamEndCapture(c->stream, &graph);
            HIPCHK(ce);
            HIPCHK(ee);
            hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (ie != hipSuccess) {
                (void)hipGraphDestroy(graph);
                HIPCHK(ie);
            }
            if (c->graphs.size() >= 8) {   // keep the most recent shapes
                (void)hipGraphExecDestroy(c->graphs.front().exec);
                (void)hipGraphDestroy(c->graphs.front().graph);
                c->graphs.erase(c->graphs.begin());
            }
            c->graphs.push_back(GraphEntry{kern, rz, br, ntiles, chunk, spl, n, k1, passes, ds, dt,
                                           base, graph, exec});
        }
        const int full = nlaunch / chunk, rem = nlaunch - full * chunk;
        HIPCHK(hipEventRecord(ev0, c->stream));   // after instantiation
        for (int k = 0; k < full; ++k) HIPCHK(hipGraphLaunch(exec, c->stream));
        for (int r = 0; r < rem; ++r) HIPCHK(step(r * spl));
    }
    HIPCHK(hipEventRecord(ev1, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
    c->t_ms += ms;
    c->t_launches += nlaunch;
    c->t_bytes += bytes;
    return MNAD_OK;
}

// Gathers of a forward shot -> column-major [nt x nr] doubles (duplicated
// receivers read the same unique point).
int download_gather(mnad_ctx* c, const Shot& s, int nt, int nr, double* out, hipStream_t st)
{
    if (!out || nr == 0) return MNAD_OK;
    if (!st) st = c->stream;
    std::vector<float> g((size_t)nt * s.nu);
    HIPCHK(hipMemcpyAsync(g.data(), s.gather.p, g.size() * sizeof(float), hipMemcpyDeviceToHost,
                          st));
    HIPCHK(hipStreamSynchronize(st));
    for (int r = 0; r < nr; ++r) {
        const int u = s.uid[r];
        for (int t = 0; t < nt; ++t) out[(size_t)t + (size_t)r * nt] = g[(size_t)t * s.nu + u];
    }
    return MNAD_OK;
}

// Final level of a field (parity nt & 1) -> column-major doubles, times scale.
int download_field(mnad_ctx* c, const Model& m, const Shot& s, int field, double scale,
                   double* out)
{
    if (!out) return MNAD_OK;
    const mnad_grid& g = m.g;
    std::vector<float> buf(m.elems);
    const int par = s.d.nt % s.d.nlev;
    HIPCHK(hipMemcpyAsync(buf.data(), s.d.f[par][field], m.elems * sizeof(float),
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < g.nx; ++i)
        for (int j = 0; j < g.nz; ++j)
            out[(size_t)j + (size_t)i * g.nz] =
                scale * buf[(size_t)(j + MNAD_ZO) * m.pitch + (i + MNAD_XO)];
    return MNAD_OK;
}

// Interior row-major float array -> column-major doubles, out = scale*(a - b).
int download_interior(mnad_ctx* c, const mnad_grid& g, const float* a, const float* b,
                      double scale, double* out, hipStream_t st)
{
    if (!out) return MNAD_OK;
    if (!st) st = c->stream;
    const int nzi = g.nz - 2 * g.npml, nxi = g.nx - 2 * g.npml;
    const size_t ni = (size_t)nzi * nxi;
    std::vector<float> ha(ni), hb(b ? ni : 0);
    HIPCHK(hipMemcpyAsync(ha.data(), a, ni * sizeof(float), hipMemcpyDeviceToHost, st));
    if (b) HIPCHK(hipMemcpyAsync(hb.data(), b, ni * sizeof(float), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int i = 0; i < nxi; ++i)
        for (int j = 0; j < nzi; ++j) {
            const size_t q = (size_t)j * nxi + i;
            out[(size_t)j + (size_t)i * nzi] = scale * ((double)ha[q] - (b ? (double)hb[q] : 0.0));
        }
    return MNAD_OK;
}

int schedule(int nt, double vmin, double vmax, int* nf, int* T, int* K, int* ncol)
{
    // fwd_PDE_exact_MNAD.m:46-47,53; Born_adj_MNAD.m:41-42.
    const double nfd = std::floor((double)nt / (2.0 * 3.0 * vmax / vmin));
    if (!(nfd >= 1.0))
        return fail(MNAD_ERR_ARG,
                    "nt = %d is too short for the snapshot schedule: nf = floor(nt/(6 vmax/vmin)) "
                    "= 0 gives T = Inf and a NaN image in the reference",
                    nt);
    const int nf_ = (int)nfd;
    const int T_ = (int)std::floor((double)nt / (double)nf_ / 2.0);
    const int K_ = nt / T_;
    if (nf) *nf = nf_;
    if (T) *T = T_;
    if (K) *K = K_;
    if (ncol) *ncol = std::max(2 * nf_ + 1, K_);
    return MNAD_OK;
}

}  // namespace mnad
