// Host-only run of the library's host code (csrc/mnad_runtime.hip,
// mnad_capi.hip, mnad_autotune.hip, mnad_server.hip) under AddressSanitizer,
// with no GPU: every HIP runtime function that code calls is replaced by a
// host fake defined here, and the kernel launchers are stubbed.
//  - device memory is host heap memory (hipMalloc = aligned_alloc), so an
//    upload, download, fill or staging copy that runs past the buffer it
//    targets is an ASan heap overflow, and a buffer used after the pool
//    freed it a use-after-free;
//  - copies and fills run at once (the fake stream is always idle); graphs,
//    events and streams are small heap objects, so a double destroy or a
//    leak of one shows too;
//  - the kernels do not run: gathers, images and fields stay as the fills
//    left them (zeros), which the host-side finiteness checks accept.
// It drives each host path twice in one process (the second call replays the
// cached graph chunks, reuses the pooled memory, the staging slots and the
// cached models): single-shot forward with kept snapshots and the adjoint,
// snapshot spill to a file and back, batched RTM (f16 snapshots, the BF16
// ring, step kernels 4 and 13), batched modelling, the autotuner, and the
// serving loop with a mix of requests.  Each step prints "ok <step>" or
// "FAIL <step>: <message>"; the exit status is the number of failures.
// Kernel-13 graph chunks: the stub of its launcher records what a capture
// bakes into a chunk (shot count, dx == dz, the passes around each launch,
// the tile count), and a replayed chunk must match the direct launches that
// follow it in the same launch sequence (the chunk cache must not hand a
// sequence a chunk captured for another batch with the same tile count).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "mnad.h"
#include "mnad_internal.h"

// ------------------------------------------------------------ HIP fakes
namespace {
struct FakeObj {
    unsigned magic = 0x5eed;
};
template <class T>
T fake_new()
{
    return reinterpret_cast<T>(new FakeObj);
}
template <class T>
hipError_t fake_delete(T p)
{
    auto* o = reinterpret_cast<FakeObj*>(p);
    if (!o || o->magic != 0x5eed) return hipErrorInvalidValue;
    o->magic = 0;
    delete o;
    return hipSuccess;
}
bool g_capturing = false;
// kernel-13 launches: what a captured chunk holds, and the chunk replayed last
struct Baked {
    int nshots = -1, k1 = -1, passes = -1, ntiles = -1;
    bool operator!=(const Baked& o) const
    {
        return nshots != o.nshots || k1 != o.k1 || passes != o.passes || ntiles != o.ntiles;
    }
};
Baked g_cap;                     // the capture in progress
std::map<void*, Baked> g_baked;  // graph or exec -> its kernel-13 launches
Baked g_pending;                 // the chunk replayed last in this sequence
bool g_pending_set = false;
int g_replays_checked = 0, g_replay_mismatches = 0;
}  // namespace

extern "C" {
hipError_t hipMalloc(void** ptr, size_t size)
{
    *ptr = std::aligned_alloc(256, (size + 255) & ~size_t(255));
    return *ptr ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* ptr)
{
    std::free(ptr);
    return hipSuccess;
}
hipError_t hipHostMalloc(void** ptr, size_t size, unsigned int)
{
    *ptr = std::malloc(size);
    return *ptr ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree(void* ptr)
{
    std::free(ptr);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t n, hipMemcpyKind, hipStream_t)
{
    if (g_capturing) return hipErrorStreamCaptureUnsupported;
    if (n) std::memcpy(dst, src, n);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void* dst, int value, size_t n, hipStream_t)
{
    if (g_capturing) return hipErrorStreamCaptureUnsupported;
    if (n) std::memset(dst, value, n);
    return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int)
{
    *s = fake_new<hipStream_t>();
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) { return fake_delete(s); }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode)
{
    g_capturing = true;
    g_cap = Baked{};
    g_pending_set = false;
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g)
{
    g_capturing = false;
    *g = fake_new<hipGraph_t>();
    g_baked[(void*)*g] = g_cap;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t)
{
    *e = fake_new<hipGraphExec_t>();
    g_baked[(void*)*e] = g_baked[(void*)g];
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t)
{
    if (reinterpret_cast<FakeObj*>(e)->magic != 0x5eed) return hipErrorInvalidValue;
    g_pending = g_baked[(void*)e];
    g_pending_set = g_pending.nshots >= 0;
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) { return fake_delete(g); }
hipError_t hipGraphExecDestroy(hipGraphExec_t e) { return fake_delete(e); }
hipError_t hipEventCreate(hipEvent_t* e)
{
    *e = fake_new<hipEvent_t>();
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) { return fake_delete(e); }
static hipError_t event_ok(hipEvent_t e)
{
    return reinterpret_cast<FakeObj*>(e)->magic == 0x5eed ? hipSuccess : hipErrorInvalidValue;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t)
{
    g_pending_set = false;   // a launch sequence ends with an event record
    return event_ok(e);
}
hipError_t hipEventSynchronize(hipEvent_t e) { return event_ok(e); }
hipError_t hipEventQuery(hipEvent_t e) { return event_ok(e); }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t)
{
    *ms = 0.25f;
    return hipSuccess;
}
hipError_t hipGetLastError(void) { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "fake HIP error"; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDeviceCount(int* n)
{
    *n = 1;
    return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int)
{
    *v = 256;
    return hipSuccess;
}
hipError_t hipMemGetInfo(size_t* freeb, size_t* total)
{
    *freeb = (size_t)4 << 30;
    *total = (size_t)8 << 30;
    return hipSuccess;
}

// ------------------------------------------------------ launcher stubs
using mnad::ResBlock;
using mnad::ShotDev;
using mnad::TileRef;
hipError_t mnad_launch_setup_model(const double*, int, int, int, int, double, double, double,
                                   float*, float*, float*, hipStream_t) { return hipSuccess; }
hipError_t mnad_launch_step(const ShotDev*, const TileRef*, int, const int*, int, int, int, int,
                            hipStream_t) { return hipSuccess; }
hipError_t mnad_launch_step2q(const ShotDev*, int nshots, const TileRef*, int ntiles, const int*,
                              int, int, int, int k1, int passes, hipStream_t)
{
    const Baked b{nshots, k1, passes, ntiles};
    if (g_capturing) {
        g_cap = b;
    } else if (g_pending_set) {   // the first direct launch after a replayed chunk
        ++g_replays_checked;
        if (g_pending != b) ++g_replay_mismatches;
        g_pending_set = false;
    }
    return hipSuccess;
}
hipError_t mnad_launch_advance(int*, int, hipStream_t) { return hipSuccess; }
hipError_t mnad_launch_init_source(ShotDev*, int, hipStream_t) { return hipSuccess; }
hipError_t mnad_launch_make_radj(const float*, const float*, const float*, int, int, float*,
                                 hipStream_t) { return hipSuccess; }
hipError_t mnad_launch_stack(const float*, const float*, const float*, float, long, float*, float*,
                             float*, float*, hipStream_t) { return hipSuccess; }
hipError_t mnad_launch_fill(float*, long, float, hipStream_t) { return hipSuccess; }
int mnad_resident_max_rows() { return 64; }
hipError_t mnad_resident_fit(int, int* per_cu)
{
    *per_cu = 2;
    return hipSuccess;
}
long mnad_resident_edge_records(int) { return 0; }
hipError_t mnad_launch_resident(const ShotDev*, const ResBlock*, int, unsigned long long*, int*,
                                int, int, int, hipStream_t) { return hipSuccess; }
}

namespace mnad {
int noalloc_init_device(int) { return MNAD_OK; }
}

// ------------------------------------------------------------- driver
static int g_bad = 0;

static void check(const char* step, int rc)
{
    if (rc == MNAD_OK) {
        printf("ok %s\n", step);
    } else {
        printf("FAIL %s: %d %s\n", step, rc, mnad_last_error());
        ++g_bad;
    }
}

int main(int argc, char** argv)
{
    const std::string dir = argc > 1 ? argv[1] : "/tmp";
    mnad_ctx* ctx = nullptr;
    check("ctx_create", mnad_ctx_create(0, &ctx));
    if (!ctx) return 1;
    // graph chunks of 8 launches, so that the 120-step runs replay graphs
    mnad_tuning t{4, 1, 1, 8, 1};
    check("set_tuning", mnad_set_tuning(ctx, &t));

    const int nz = 48, nx = 64, npml = 8, nt = 120;
    const mnad_grid g{nz, nx, npml, 10.0, 10.0, 1e-3};
    std::vector<double> v((size_t)nz * nx), vs((size_t)nz * nx), w(nt);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < nz; ++j) {
            v[(size_t)j + (size_t)i * nz] = j < nz / 2 ? 1800.0 : 2400.0;
            vs[(size_t)j + (size_t)i * nz] = 1800.0 + 600.0 * j / nz;
        }
    for (int k = 0; k < nt; ++k) w[k] = (k == 5) ? 1.0 : 0.0;
    const int nr = nx - 2 * npml;
    std::vector<int> zr(nr, npml + 1), xr(nr);
    for (int r = 0; r < nr; ++r) xr[r] = npml + r;
    const int nzi = nz - 2 * npml, nxi = nx - 2 * npml;
    std::vector<double> data((size_t)nt * nr), img((size_t)nzi * nxi), ill((size_t)nzi * nxi);
    std::vector<double> U((size_t)nz * nx), V((size_t)nz * nx), W((size_t)nz * nx);

    for (int rep = 0; rep < 2; ++rep) {
        // single shot: forward keeping snapshots, spill and reload, adjoint
        check("fwd_keep", mnad_fwd(ctx, &g, vs.data(), nt, w.data(), npml, nx / 2, nr, zr.data(),
                                   xr.data(), 3, 1, MNAD_SNAP_F16, 0, data.data(), U.data(),
                                   V.data(), W.data(), ill.data()));
        const std::string path = dir + "/snaps_" + std::to_string(rep) + ".bin";
        check("snap_save", mnad_snap_save(ctx, 3, path.c_str()));
        check("snap_load", mnad_snap_load(ctx, 3, path.c_str()));
        check("born_adj", mnad_born_adj(ctx, &g, vs.data(), nt, nr, zr.data(), xr.data(),
                                        data.data(), 3, 0, img.data()));
        check("fwd_plain", mnad_fwd(ctx, &g, v.data(), nt, w.data(), npml, nx / 2, nr, zr.data(),
                                    xr.data(), 0, 0, MNAD_SNAP_F32, MNAD_FLAG_NO_GRAPH,
                                    data.data(), nullptr, nullptr, nullptr, nullptr));
        // batched RTM of three shots: f16 snapshots, then the BF16 ring, two per group
        const int ns = 3;
        std::vector<int> zs(ns, npml), xs = {npml + 2, nx / 2, nx - npml - 3}, off(ns + 1);
        std::vector<int> azr, axr;
        for (int k = 0; k < ns; ++k) {
            off[k] = (int)azr.size();
            for (int r = 0; r < nr; r += 1 + k) {
                azr.push_back(zr[r]);
                axr.push_back(xr[r]);
            }
        }
        off[ns] = (int)azr.size();
        const mnad_shots sh{ns, zs.data(), xs.data(), off.data(), azr.data(), axr.data()};
        std::vector<double> dobs((size_t)nt * azr.size()), dsyn((size_t)nt * azr.size());
        for (int fmt : {MNAD_SNAP_F16, MNAD_SNAP_RING_BF16}) {
            int sid = -1;
            check("stack_create", mnad_stack_create(ctx, &g, &sid));
            check(fmt == MNAD_SNAP_F16 ? "rtm_shots_f16" : "rtm_shots_ring",
                  mnad_rtm_shots(ctx, sid, &g, v.data(), vs.data(), nt, w.data(), &sh, fmt, 2, 0,
                                 dobs.data(), dsyn.data()));
            int n = 0;
            check("stack_image", mnad_stack_image(ctx, sid, img.data(), ill.data(), &n));
            check("stack_destroy", mnad_stack_destroy(ctx, sid));
        }
        // the same on the two-step kernel 13 (BF16 ring)
        mnad_tuning t13{13, 2, 4, 8, 1};
        check("set_tuning_13", mnad_set_tuning(ctx, &t13));
        {
            int sid = -1;
            check("stack_create_13", mnad_stack_create(ctx, &g, &sid));
            check("rtm_shots_ring_13", mnad_rtm_shots(ctx, sid, &g, v.data(), vs.data(), nt,
                                                      w.data(), &sh, MNAD_SNAP_RING_BF16, 3, 0,
                                                      nullptr, nullptr));
            check("stack_destroy_13", mnad_stack_destroy(ctx, sid));
        }
        {   // kernel 13 on a 241 x 241 grid (npml 20), BF16 ring, two shots in one
            // group (the second grid of the GPU ring tests): the forward sequence
            // and the sequence of adjoints and reconstructions replay chunks
            const int n1 = 241, p1 = 20, nt1 = 122;
            const mnad_grid g1{n1, n1, p1, 10.0, 10.0, 1e-3};
            std::vector<double> v1((size_t)n1 * n1), w1(nt1, 0.0);
            for (int i = 0; i < n1; ++i)
                for (int j = 0; j < n1; ++j) v1[(size_t)j + (size_t)i * n1] = j < 120 ? 1500.0 : 2500.0;
            w1[5] = 1.0;
            std::vector<int> z1(2, p1), x1 = {p1 + 100, p1 + 137}, o1 = {0, 201, 402}, zr1, xr1;
            for (int k = 0; k < 2; ++k)
                for (int r = 0; r < 201; ++r) {
                    zr1.push_back(p1 + 1);
                    xr1.push_back(p1 + r);
                }
            const mnad_shots sh1{2, z1.data(), x1.data(), o1.data(), zr1.data(), xr1.data()};
            int sid = -1;
            check("stack_create_13_c1", mnad_stack_create(ctx, &g1, &sid));
            check("rtm_shots_ring_13_c1", mnad_rtm_shots(ctx, sid, &g1, v1.data(), v1.data(), nt1,
                                                         w1.data(), &sh1, MNAD_SNAP_RING_BF16, 2,
                                                         0, nullptr, nullptr));
            check("stack_destroy_13_c1", mnad_stack_destroy(ctx, sid));
        }
        check("set_tuning_4", mnad_set_tuning(ctx, &t));
        check("fwd_shots", mnad_fwd_shots(ctx, &g, v.data(), nt, w.data(), &sh, 2, 0, dobs.data()));
    }
    {
        mnad_tuning best{};
        int shots = 0;
        double us = 0;
        check("autotune", mnad_autotune(ctx, &g, v.data(), 4, &best, &shots, &us));
        check("set_tuning_back", mnad_set_tuning(ctx, &t));
    }
    {   // serving: two modelling and two RTM requests, in two rounds
        check("serve_start", mnad_serve_start(ctx, 0, 8, 0, 16));
        for (int round = 0; round < 2; ++round) {
            std::vector<std::vector<double>> outs(4), imgs(4);
            std::vector<int64_t> tk(4, -1);
            for (int k = 0; k < 4; ++k) {
                mnad_request q{};
                q.kind = k % 2 ? MNAD_REQ_RTM : MNAD_REQ_MODEL;
                q.grid = g;
                q.v = v.data();
                q.v_smooth = vs.data();
                q.nt = nt - 8 * k;
                q.w = w.data();
                q.zs = npml;
                q.xs = npml + 4 + 7 * k;
                q.nr = nr;
                q.zr = zr.data();
                q.xr = xr.data();
                q.snap_fmt = MNAD_SNAP_F16;
                outs[k].assign((size_t)q.nt * nr, 0.0);
                imgs[k].assign((size_t)nzi * nxi, 0.0);
                q.out_gather = outs[k].data();
                q.out_image = q.kind == MNAD_REQ_RTM ? imgs[k].data() : nullptr;
                check("submit", mnad_submit(ctx, &q, &tk[k]));
            }
            for (int k = 0; k < 4; ++k) {
                mnad_req_info info{};
                check("wait", mnad_wait(ctx, tk[k], -1, &info));
            }
        }
        check("serve_stop", mnad_serve_stop(ctx));
    }
    mnad_ctx_destroy(ctx);
    if (g_replays_checked == 0 || g_replay_mismatches > 0) {
        printf("FAIL graph_chunk_replays: %d of %d kernel-13 replays do not match their sequence\n",
               g_replay_mismatches, g_replays_checked);
        ++g_bad;
    } else {
        printf("ok graph_chunk_replays (%d)\n", g_replays_checked);
    }
    printf("%d failure(s)\n", g_bad);
    return g_bad;
}
