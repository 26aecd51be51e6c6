// mnad_server.hip -- serving loop with continuous batching (mnad_serve_*,
// mnad_submit / mnad_wait / mnad_cancel).
//
// The reference runs one shot per parfor worker (main_Acoustic_Data.m:169-171,
// main_Acoustic_RTM.m:223-226,250-254) and has no serving path.  Here one
// server thread keeps every admitted request's wavefields resident and steps
// all of their shots together: each shot carries its own origin t0 on the
// global step counter, so requests that arrive while others run join the next
// launch without waiting for a batch to drain, and leave it when their last
// step has run.  A shot's arithmetic does not depend on its neighbours in the
// launch, so each result equals the one computed alone.
//
// The loop keeps two chunks of steps on the stream: while the newer one runs,
// it waits for the older one, downloads the results it finished on a second
// stream, admits new requests (their set-up is queued behind the running
// chunk) and queues the next chunk.  A chunk is a captured graph of its step
// launches, replayed while the set of active shots keeps the same (padded)
// tile count.  Device memory a finished or cancelled request gives back is
// reused only by work queued after the chunks that may still touch it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "mnad.h"
#include "mnad_internal.h"

namespace mnad {

extern "C" {
hipError_t mnad_launch_step(const ShotDev* shots, const TileRef* tiles, int ntiles,
                            const int* base, int rel, int kernel, int rows, int brows,
                            hipStream_t st);
hipError_t mnad_launch_advance(int* base, int n, hipStream_t st);
hipError_t mnad_launch_init_source(ShotDev* shots, int nshots, hipStream_t st);
hipError_t mnad_launch_make_radj(const float* dobs, const float* dsyn, const float* arec,
                                 int nt, int nrec, float* radj, hipStream_t st);
}

namespace {

enum ReqState { QUEUED, ACTIVE, DONE };

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
    int64_t id = 0;
    mnad_request spec{};
    std::vector<double> v, vs, w;
    std::vector<int> zr, xr;
    ReqState state = QUEUED;
    int rc = MNAD_OK;
    std::string err;
    bool cancel = false;
    bool claimed = false;   // a mnad_wait has taken the result
    double t_submit = 0, t_admit = 0, t_done = 0;
    size_t bytes = 0;
    size_t points = 0;   // grid points stepping while the request is active
    // device side
    std::unique_ptr<Model> mt, ms;
    std::vector<std::unique_ptr<Shot>> shots;   // the shots currently stepping
    std::unique_ptr<Shot> keep_syn;             // RTM: migration-model forward (snapshots, illum)
    int phase = 0;                              // RTM: 0 forwards, 1 adjoint
    int T = 0, K = 0, ncol = 0;
};

size_t field_bytes(const mnad_grid& g)
{
    const size_t pitch = ((MNAD_XO + g.nx + 2) + 31) / 32 * 32;
    return (size_t)(g.nz + 2 * MNAD_ZO + MNAD_ZPAD) * pitch * 4;
}

}  // namespace

struct Server {
    mnad_ctx* c = nullptr;   // private context: own stream and pool
    hipStream_t copy = nullptr;   // downloads of finished requests, beside the next chunk
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queued;
    std::vector<Request*> active;
    // every request not yet waited for; a waiter holds its own reference
    std::map<int64_t, std::shared_ptr<Request>> all;
    size_t budget = 0, reserved = 0;
    size_t max_points = 0, active_points = 0;
    int max_active = 64, chunk = 32;
    bool stop = false;
    int64_t next_id = 1;
    // Global step counter: G steps are enqueued (the device counter *dbase
    // reaches G once they have run), G_done of them are known to have run.
    long long G = 0, G_done = 0;
    DevBuf dsh, dtab, dbase;   // descriptors and tile table of the active shots (grow-only)
    int ntiles = 0;
    bool changed = true;
    // Chunks on the stream not yet waited for (at most kInflight).
    struct Inflight {
        long long g_end;
        hipEvent_t ev;
    };
    std::deque<Inflight> inflight;
    std::vector<hipEvent_t> spare;   // events of finished chunks, for reuse
    // Captured chunks: `chunk` step launches and the counter advance.
    struct ChunkGraph {
        int kernel, rz, brows, ntiles;
        const void* ds;
        const void* dt;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<ChunkGraph> graphs;

    void run();
    int admit(Request* r);
    int start_adjoint(Request* r);
    void complete(Request* r, int rc, const char* msg);
    int enqueue_chunk();
    int wait_oldest();
    int drain();
    void drop_graphs(const void* ptr);
    void finish_ready();
};

constexpr int kInflight = 2;

// Tile count a chunk graph is captured for: the table is padded (entries that
// exit at once) to m 2^e with m in 4 .. 8, so that the graphs of a changing
// mix of requests repeat.
static size_t padded_tiles(size_t n)
{
    if (n <= 8) return 8;
    size_t p = 1;
    while (p * 8 < n) p *= 2;
    size_t m = 4;
    while (m * p < n) ++m;
    return (m * p + 7) / 8 * 8;
}

// Conservative device bytes of a request at its peak.
static size_t request_bytes(const mnad_request& q)
{
    const mnad_grid& g = q.grid;
    const size_t np = g.npml;
    const size_t model = field_bytes(g) + 3 * 4 * ((size_t)g.nz * 2 * np + 2 * np * g.nx) +
                         (size_t)g.nz * g.nx * 8;
    const size_t shot = 6 * field_bytes(g) + 9 * 4 * ((size_t)g.nz * 2 * np + 2 * np * g.nx) +
                        (size_t)q.nt * q.nr * 4 + (size_t)(q.nt + 1) * 4 + (size_t)g.nz * 4 +
                        (size_t)g.nx * 4 * 4 + 4096;
    if (q.kind == MNAD_REQ_MODEL) return model + shot;
    size_t ni = (size_t)(g.nz - 2 * np) * (g.nx - 2 * np);
    double vmin = q.v_smooth[0], vmax = q.v_smooth[0];
    for (size_t k = 0; k < (size_t)g.nz * g.nx; ++k) {
        vmin = std::min(vmin, q.v_smooth[k]);
        vmax = std::max(vmax, q.v_smooth[k]);
    }
    int T = 0, K = 0, ncol = 0;
    if (schedule(q.nt, vmin, vmax, nullptr, &T, &K, &ncol) != MNAD_OK) K = 0;
    const size_t esz = q.snap_fmt == MNAD_SNAP_F32 ? 4 : 2;
    return 2 * model + 3 * shot + (size_t)K * ni * esz + 4 * ni * 4;
}

void Server::complete(Request* r, int rc, const char* msg)
{
    r->shots.clear();
    r->keep_syn.reset();
    r->mt.reset();
    r->ms.reset();
    std::lock_guard<std::mutex> lk(mu);
    r->rc = rc;
    if (msg) r->err = msg;
    r->state = DONE;
    r->t_done = now_s();
    reserved -= std::min(reserved, r->bytes);
    if (r->t_admit > 0) active_points -= std::min(active_points, r->points);
    cv_done.notify_all();
}

// Device set-up of an admitted request (server thread, outside the lock).
int Server::admit(Request* r)
{
    const mnad_request& q = r->spec;
    r->mt = std::make_unique<Model>();
    int rc = make_model(c, &q.grid, r->v.data(), *r->mt);
    if (rc != MNAD_OK) return rc;
    ShotSpec sp;
    sp.mode = MODE_FWD;
    sp.nt = q.nt;
    sp.w = r->w.data();
    sp.zs = q.zs;
    sp.xs = q.xs;
    sp.nr = q.nr;
    sp.zr = r->zr.data();
    sp.xr = r->xr.data();
    auto obs = std::make_unique<Shot>();
    rc = make_shot(c, *r->mt, sp, *obs);
    if (rc != MNAD_OK) return rc;
    obs->t0 = G;
    r->shots.push_back(std::move(obs));
    if (q.kind == MNAD_REQ_RTM) {
        r->ms = std::make_unique<Model>();
        rc = make_model(c, &q.grid, r->vs.data(), *r->ms);
        if (rc != MNAD_OK) return rc;
        rc = schedule(q.nt, r->ms->vmin, r->ms->vmax, nullptr, &r->T, &r->K, &r->ncol);
        if (rc != MNAD_OK) return rc;
        sp.T = r->T;
        sp.K = r->K;
        sp.ncol = r->ncol;
        sp.snap_fmt = q.snap_fmt;
        auto syn = std::make_unique<Shot>();
        rc = make_shot(c, *r->ms, sp, *syn);
        if (rc != MNAD_OK) return rc;
        syn->t0 = G;
        r->shots.push_back(std::move(syn));
    }
    // first injection into the new shots' initial U
    std::vector<ShotDev> h;
    for (auto& s : r->shots) h.push_back(s->d);
    DevBuf tmp;
    rc = upload(c, tmp, h.data(), h.size() * sizeof(ShotDev));
    if (rc != MNAD_OK) return rc;
    // enqueued behind the chunks in flight, which do not touch these shots;
    // tmp's memory goes back to the pool, and any later use of it is ordered
    // after this launch on the same stream
    if (mnad_launch_init_source(tmp.as<ShotDev>(), (int)h.size(), c->stream) != hipSuccess)
        return fail(MNAD_ERR_HIP, "init_source launch failed");
    return MNAD_OK;
}

// RTM: both forwards done -> adjoint with imaging on the migration model.
int Server::start_adjoint(Request* r)
{
    const mnad_request& q = r->spec;
    Shot& obs = *r->shots[0];
    Shot& syn = *r->shots[1];
    if (q.out_gather) {   // obs has run its last step: read it beside the next chunk
        int rc = download_gather(c, obs, q.nt, q.nr, q.out_gather, copy);
        if (rc != MNAD_OK) return rc;
    }
    ShotSpec sp;
    sp.mode = MODE_ADJ;
    sp.nt = q.nt;
    sp.nr = q.nr;
    sp.zr = r->zr.data();
    sp.xr = r->xr.data();
    sp.T = r->T;
    sp.K = r->K;
    sp.ncol = r->ncol;
    auto adj = std::make_unique<Shot>();
    int rc = make_shot(c, *r->ms, sp, *adj);
    if (rc != MNAD_OK) return rc;
    if (adj->nu > 0 &&
        mnad_launch_make_radj(obs.gather.as<float>(), syn.gather.as<float>(),
                              adj->arec.as<float>(), q.nt, adj->nu, adj->radj.as<float>(),
                              c->stream) != hipSuccess)
        return fail(MNAD_ERR_HIP, "make_radj launch failed");
    attach_snaps(*adj, std::move(syn.snaps), q.snap_fmt, syn.d.sscale);
    adj->t0 = G;
    syn.release_state();
    r->keep_syn = std::move(r->shots[1]);   // keeps the illumination buffer
    r->shots.clear();                       // observed forward released
    r->shots.push_back(std::move(adj));
    r->phase = 1;
    return MNAD_OK;
}

void Server::drop_graphs(const void* ptr)
{
    for (size_t k = 0; k < graphs.size();) {
        if (!ptr || graphs[k].ds == ptr || graphs[k].dt == ptr) {
            (void)hipGraphExecDestroy(graphs[k].exec);
            (void)hipGraphDestroy(graphs[k].graph);
            graphs.erase(graphs.begin() + (long)k);
        } else {
            ++k;
        }
    }
}

// Waits for the oldest chunk in flight; G_done moves to its end.
int Server::wait_oldest()
{
    if (inflight.empty()) return MNAD_OK;
    const Inflight f = inflight.front();
    inflight.pop_front();
    spare.push_back(f.ev);
    if (hipEventSynchronize(f.ev) != hipSuccess) return fail(MNAD_ERR_HIP, "step chunk failed");
    G_done = f.g_end;
    return MNAD_OK;
}

int Server::drain()
{
    while (!inflight.empty()) {
        const int rc = wait_oldest();
        if (rc != MNAD_OK) return rc;
    }
    return MNAD_OK;
}

// Enqueues one chunk of `chunk` steps over every active shot, without
// waiting for the chunks before it.  A shot whose steps are all done exits
// the kernel at its first check (its descriptor may still be listed until
// the host has seen it finish).
int Server::enqueue_chunk()
{
    if (G > (1LL << 30)) {   // keep the device counter far from int overflow
        int rc = drain();
        if (rc != MNAD_OK) return rc;
        for (Request* r : active)
            for (auto& s : r->shots) s->t0 -= G;
        G = 0;
        G_done = 0;
        if (hipMemsetAsync(dbase.p, 0, sizeof(int), c->stream) != hipSuccess)
            return fail(MNAD_ERR_HIP, "reset of the step counter failed");
        changed = true;
    }
    if (changed) {
        std::vector<ShotDev> h;
        for (Request* r : active)
            for (auto& s : r->shots) {
                ShotDev d = s->d;
                d.t0 = (int)s->t0;
                h.push_back(d);
            }
        if (h.empty()) return MNAD_OK;
        std::vector<TileRef> table;
        build_tiles(c, h, table);
        table.resize(padded_tiles(table.size()), TileRef{-1, 0});
        const size_t bh = h.size() * sizeof(ShotDev), bt = table.size() * sizeof(TileRef);
        if (dsh.n < bh) {   // grow-only, so that captured chunks stay valid
            drop_graphs(dsh.p);
            int rc = dalloc(c, dsh, std::max(bh, 2 * dsh.n), false);
            if (rc != MNAD_OK) return rc;
        }
        if (dtab.n < bt) {
            drop_graphs(dtab.p);
            int rc = dalloc(c, dtab, std::max(bt, 2 * dtab.n), false);
            if (rc != MNAD_OK) return rc;
        }
        int rc = copy_in(c, dsh.p, h.data(), bh);
        if (rc == MNAD_OK) rc = copy_in(c, dtab.p, table.data(), bt);
        if (rc != MNAD_OK) return rc;
        ntiles = (int)table.size();
        changed = false;
    }
    const ShotDev* ds = dsh.as<ShotDev>();
    const TileRef* dt = dtab.as<TileRef>();
    int* base = dbase.as<int>();
    hipGraphExec_t exec = nullptr;
    for (const ChunkGraph& e : graphs)
        if (e.kernel == c->kernel && e.rz == c->rz && e.brows == c->brows && e.ntiles == ntiles &&
            e.ds == ds && e.dt == dt) {
            exec = e.exec;
            break;
        }
    if (!exec) {
        hipGraph_t graph = nullptr;
        if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return fail(MNAD_ERR_HIP, "chunk capture failed to start");
        hipError_t ce = hipSuccess;
        for (int r = 0; r < chunk && ce == hipSuccess; ++r)
            ce = mnad_launch_step(ds, dt, ntiles, base, r, c->kernel, c->rz, c->brows, c->stream);
        if (ce == hipSuccess) ce = mnad_launch_advance(base, chunk, c->stream);
        const hipError_t ee = hipStreamEndCapture(c->stream, &graph);
        if (ce != hipSuccess || ee != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return fail(MNAD_ERR_HIP, "chunk capture failed");
        }
        if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return fail(MNAD_ERR_HIP, "chunk graph instantiation failed");
        }
        if (graphs.size() >= 16) {   // keep the most recent shapes
            (void)hipGraphExecDestroy(graphs.front().exec);
            (void)hipGraphDestroy(graphs.front().graph);
            graphs.erase(graphs.begin());
        }
        graphs.push_back(ChunkGraph{c->kernel, c->rz, c->brows, ntiles, ds, dt, graph, exec});
    }
    if (hipGraphLaunch(exec, c->stream) != hipSuccess) return fail(MNAD_ERR_HIP, "chunk launch failed");
    hipEvent_t ev = nullptr;
    if (!spare.empty()) {
        ev = spare.back();
        spare.pop_back();
    } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        return fail(MNAD_ERR_HIP, "event creation failed");
    }
    if (hipEventRecord(ev, c->stream) != hipSuccess) {
        spare.push_back(ev);
        return fail(MNAD_ERR_HIP, "event record failed");
    }
    G += chunk;
    inflight.push_back(Inflight{G, ev});
    return MNAD_OK;
}

// Requests whose shots have all run their last step move on or complete.
void Server::finish_ready()
{
    std::vector<Request*> still;
    for (Request* r : active) {
        bool done = !r->shots.empty();
        for (auto& s : r->shots) done = done && (G_done - s->t0 >= s->d.nt);
        if (!done) {
            still.push_back(r);
            continue;
        }
        const mnad_request& q = r->spec;
        int rc = MNAD_OK;
        if (q.kind == MNAD_REQ_MODEL) {
            rc = download_gather(c, *r->shots[0], q.nt, q.nr, q.out_gather, copy);
            if (rc == MNAD_OK && q.out_gather)
                for (size_t k = 0; k < (size_t)q.nt * q.nr; ++k)
                    if (!std::isfinite(q.out_gather[k])) {
                        rc = fail(MNAD_ERR_ARG, "gather is not finite (unstable request)");
                        break;
                    }
            complete(r, rc, rc == MNAD_OK ? nullptr : last_error());
            changed = true;
        } else if (r->phase == 0) {
            rc = start_adjoint(r);
            changed = true;
            if (rc != MNAD_OK) complete(r, rc, last_error());
            else still.push_back(r);
        } else {
            const mnad_grid& g = q.grid;
            Shot& adj = *r->shots[0];
            rc = download_interior(c, g, adj.img.as<float>(), adj.imgc.as<float>(), (double)r->T,
                                   q.out_image, copy);
            if (rc == MNAD_OK && q.out_illum)
                rc = download_interior(c, g, r->keep_syn->illum.as<float>(), nullptr,
                                       (double)r->T, q.out_illum, copy);
            if (rc == MNAD_OK && q.out_image) {
                const size_t ni = (size_t)(g.nz - 2 * g.npml) * (g.nx - 2 * g.npml);
                for (size_t k = 0; k < ni; ++k)
                    if (!std::isfinite(q.out_image[k])) {
                        rc = fail(MNAD_ERR_ARG, "image is not finite (unstable request)");
                        break;
                    }
            }
            complete(r, rc, rc == MNAD_OK ? nullptr : last_error());
            changed = true;
        }
    }
    active.swap(still);
}

void Server::run()
{
    (void)hipSetDevice(c->device);
    for (;;) {
        std::vector<Request*> admitted;
        {
            std::unique_lock<std::mutex> lk(mu);
            if (active.empty() && queued.empty() && !stop) {
                // idle: let the chunks still on the stream finish first
                lk.unlock();
                (void)drain();
                lk.lock();
            }
            cv_work.wait(lk, [&] { return stop || !queued.empty() || !active.empty(); });
            if (stop) break;
            // cancellations of active requests take effect at this boundary
            for (Request* r : active)
                if (r->cancel) r->rc = MNAD_ERR_CANCELLED;
            // admission: first come, first served, while memory, slots and
            // the work limit allow (the first request of an idle server always fits)
            while (!queued.empty()) {
                Request* r = queued.front();
                if ((int)(active.size() + admitted.size()) >= max_active) break;
                if (reserved + r->bytes > budget) break;
                if (max_points && active_points > 0 && active_points + r->points > max_points)
                    break;
                reserved += r->bytes;
                active_points += r->points;
                r->state = ACTIVE;
                r->t_admit = now_s();
                admitted.push_back(r);
                queued.pop_front();
            }
        }
        // drop cancelled active requests
        {
            std::vector<Request*> still;
            for (Request* r : active) {
                if (r->rc == MNAD_ERR_CANCELLED) {
                    complete(r, MNAD_ERR_CANCELLED, "cancelled");
                    changed = true;
                } else {
                    still.push_back(r);
                }
            }
            active.swap(still);
        }
        for (Request* r : admitted) {
            const int rc = admit(r);
            if (rc != MNAD_OK) {
                complete(r, rc, last_error());
            } else {
                active.push_back(r);
                changed = true;
            }
        }
        if (active.empty()) continue;
        // one chunk more on the stream; with kInflight chunks there, wait for
        // the oldest and move on the requests it finished, while the newer
        // chunk keeps the device busy
        int rc = enqueue_chunk();
        if (rc == MNAD_OK && (int)inflight.size() >= kInflight) rc = wait_oldest();
        if (rc != MNAD_OK) {   // device error: fail every active request, keep serving
            const std::string msg = last_error();
            while (!inflight.empty()) {
                (void)hipEventSynchronize(inflight.front().ev);
                spare.push_back(inflight.front().ev);
                inflight.pop_front();
            }
            (void)hipGetLastError();
            G_done = G;
            for (Request* r : active) complete(r, rc, msg.c_str());
            active.clear();
            changed = true;
            continue;
        }
        finish_ready();
    }
    // shutdown: everything still queued or active is cancelled
    (void)drain();
    for (Request* r : active) complete(r, MNAD_ERR_CANCELLED, "server stopped");
    active.clear();
    std::lock_guard<std::mutex> lk(mu);
    for (Request* r : queued) {
        r->rc = MNAD_ERR_CANCELLED;
        r->err = "server stopped";
        r->state = DONE;
        r->t_done = now_s();
    }
    queued.clear();
    cv_done.notify_all();
}

// The context's server, with a reference that keeps it alive for the caller.
static std::shared_ptr<Server> get_server(mnad_ctx* ctx)
{
    std::lock_guard<std::mutex> lk(ctx->srv_mu);
    return ctx->server;
}

// Stops the loop: new calls no longer find the server, the thread cancels
// everything still queued or active and wakes every waiter, and the device
// side goes away.  The Server object itself (with the requests not yet waited
// for) is freed when the last caller still inside mnad_wait & co. returns.
void server_destroy(mnad_ctx* ctx)
{
    std::shared_ptr<Server> s;
    {
        std::lock_guard<std::mutex> lk(ctx->srv_mu);
        s.swap(ctx->server);
    }
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->th.joinable()) s->th.join();
    (void)hipSetDevice(s->c->device);
    (void)hipStreamSynchronize(s->c->stream);
    s->drop_graphs(nullptr);
    for (const Server::Inflight& f : s->inflight) (void)hipEventDestroy(f.ev);
    s->inflight.clear();
    for (hipEvent_t e : s->spare) (void)hipEventDestroy(e);
    s->spare.clear();
    if (s->copy) (void)hipStreamDestroy(s->copy);
    s->copy = nullptr;
    s->dsh.reset();
    s->dtab.reset();
    s->dbase.reset();
    mnad_ctx_destroy(s->c);
    s->c = nullptr;
}

}  // namespace mnad

using namespace mnad;

extern "C" int mnad_serve_start(mnad_ctx* ctx, uint64_t hbm_budget, int max_active,
                                uint64_t max_points, int chunk_steps)
{
    if (!ctx) return fail(MNAD_ERR_ARG, "context is NULL");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (get_server(ctx)) return fail(MNAD_ERR_STATE, "the server is already running");
    auto s = std::make_shared<Server>();
    int rc = mnad_ctx_create(ctx->device, &s->c);
    if (rc != MNAD_OK) return rc;
    // The server steps one time level per launch (admission at any step), so
    // the two-step and the whole-batch kernels fall back to the default.  In
    // saturated C4 mixes the default row kernel (4) ran 102-104 Gpt-updates/s
    // and the four-column kernel (6) 89, also with more points admitted.
    const bool other = two_step_kernel(ctx->kernel) || resident_kernel(ctx->kernel);
    s->c->kernel = other ? kDefaultKernel : ctx->kernel;
    s->c->rz = other ? kDefaultRows : ctx->rz;
    s->c->brows = other ? kDefaultBrows : ctx->brows;
    s->c->strip_first = ctx->strip_first;
    s->c->strip_weight = ctx->strip_weight;
    s->c->x4_weight = ctx->x4_weight;
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) (void)hipGetLastError();
    s->budget = hbm_budget ? (size_t)hbm_budget : (size_t)(0.8 * (double)freeb);
    s->max_active = max_active > 0 ? max_active : 64;
    s->max_points = (size_t)max_points;
    s->chunk = chunk_steps > 0 ? chunk_steps : 32;
    rc = dalloc(s->c, s->dbase, sizeof(int), true);
    if (rc == MNAD_OK && hipStreamSynchronize(s->c->stream) != hipSuccess)
        rc = fail(MNAD_ERR_HIP, "sync failed");
    if (rc == MNAD_OK && hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking) != hipSuccess)
        rc = fail(MNAD_ERR_HIP, "copy stream creation failed");
    if (rc != MNAD_OK) {
        s->dbase.reset();
        mnad_ctx_destroy(s->c);
        return rc;
    }
    Server* raw = s.get();
    s->th = std::thread([raw] { raw->run(); });
    std::lock_guard<std::mutex> slk(ctx->srv_mu);
    ctx->server = std::move(s);
    return MNAD_OK;
}

extern "C" int mnad_serve_stop(mnad_ctx* ctx)
{
    if (!ctx) return fail(MNAD_ERR_ARG, "context is NULL");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    server_destroy(ctx);
    return MNAD_OK;
}

extern "C" int mnad_submit(mnad_ctx* ctx, const mnad_request* q, int64_t* ticket)
{
    if (!ctx || !q || !ticket) return fail(MNAD_ERR_ARG, "NULL argument");
    std::shared_ptr<Server> s = get_server(ctx);
    if (!s) return fail(MNAD_ERR_STATE, "the server is not running (mnad_serve_start)");
    if (q->kind != MNAD_REQ_MODEL && q->kind != MNAD_REQ_RTM)
        return fail(MNAD_ERR_ARG, "unknown request kind %d", q->kind);
    const mnad_grid& g = q->grid;
    if (g.nz < 3 || g.nx < 3 || g.npml < 0 || g.nz < 2 * g.npml + 1 || g.nx < 2 * g.npml + 1)
        return fail(MNAD_ERR_ARG, "bad grid %dx%d (npml %d)", g.nz, g.nx, g.npml);
    if (!q->v || !q->w || q->nt < 1 || q->nr < 0 || (q->nr > 0 && (!q->zr || !q->xr)))
        return fail(MNAD_ERR_ARG, "missing velocity, wavelet or receivers");
    if (q->kind == MNAD_REQ_RTM && !q->v_smooth)
        return fail(MNAD_ERR_ARG, "an RTM request needs v_smooth");
    if (q->kind == MNAD_REQ_MODEL && !q->out_gather)
        return fail(MNAD_ERR_ARG, "a modelling request needs out_gather");
    if (q->kind == MNAD_REQ_RTM && !q->out_image)
        return fail(MNAD_ERR_ARG, "an RTM request needs out_image");
    if (q->snap_fmt < MNAD_SNAP_F32 || q->snap_fmt > MNAD_SNAP_BF16)
        return fail(MNAD_ERR_ARG, "unknown snapshot format %d", q->snap_fmt);
    const size_t n = (size_t)g.nz * g.nx;
    // Validation the device set-up would otherwise reject later (NaN, CFL, geometry).
    for (int pass = 0; pass < (q->kind == MNAD_REQ_RTM ? 2 : 1); ++pass) {
        const double* v = pass ? q->v_smooth : q->v;
        double vmax = 0;
        for (size_t k = 0; k < n; ++k) {
            if (!std::isfinite(v[k]) || !(v[k] > 0))
                return fail(MNAD_ERR_ARG, "velocity must be finite and positive");
            vmax = std::max(vmax, v[k]);
        }
        if (vmax * g.dt / std::min(g.dx, g.dz) > 1.0)
            return fail(MNAD_ERR_ARG, "vmax*dt/min(dx,dz) > 1: unstable request");
    }
    if (q->zs < 0 || q->zs >= g.nz || q->xs < 0 || q->xs >= g.nx)
        return fail(MNAD_ERR_ARG, "source outside the grid");
    for (int r = 0; r < q->nr; ++r)
        if (q->zr[r] < 0 || q->zr[r] >= g.nz || q->xr[r] < 0 || q->xr[r] >= g.nx)
            return fail(MNAD_ERR_ARG, "receiver %d outside the grid", r);
    auto r = std::make_shared<Request>();
    r->spec = *q;
    r->v.assign(q->v, q->v + n);
    if (q->kind == MNAD_REQ_RTM) r->vs.assign(q->v_smooth, q->v_smooth + n);
    r->w.assign(q->w, q->w + q->nt);
    r->zr.assign(q->zr, q->zr + q->nr);
    r->xr.assign(q->xr, q->xr + q->nr);
    r->spec.v = r->v.data();
    r->spec.v_smooth = q->kind == MNAD_REQ_RTM ? r->vs.data() : nullptr;
    r->spec.w = r->w.data();
    r->spec.zr = r->zr.data();
    r->spec.xr = r->xr.data();
    if (q->kind == MNAD_REQ_RTM) {
        double vmin = r->vs[0], vmax = r->vs[0];
        for (double x : r->vs) {
            vmin = std::min(vmin, x);
            vmax = std::max(vmax, x);
        }
        int rc = schedule(q->nt, vmin, vmax, nullptr, nullptr, nullptr, nullptr);
        if (rc != MNAD_OK) return rc;
    }
    r->bytes = request_bytes(r->spec);
    r->points = (size_t)g.nz * g.nx * (q->kind == MNAD_REQ_RTM ? 2 : 1);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return fail(MNAD_ERR_STATE, "the server is stopping");
    if (r->bytes > s->budget)
        return fail(MNAD_ERR_BUDGET, "request needs %zu bytes, the HBM budget is %zu", r->bytes,
                    s->budget);
    r->id = s->next_id++;
    r->t_submit = now_s();
    *ticket = r->id;
    s->queued.push_back(r.get());
    s->all.emplace(r->id, std::move(r));
    s->cv_work.notify_all();
    return MNAD_OK;
}

extern "C" int mnad_wait(mnad_ctx* ctx, int64_t ticket, int timeout_ms, mnad_req_info* info)
{
    if (!ctx) return fail(MNAD_ERR_ARG, "context is NULL");
    std::shared_ptr<Server> s = get_server(ctx);
    if (!s) return fail(MNAD_ERR_STATE, "the server is not running");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return fail(MNAD_ERR_STATE, "unknown ticket %lld", (long long)ticket);
    std::shared_ptr<Request> r = it->second;   // stays valid whatever another waiter does
    // a stopping server marks every request done before it wakes the waiters
    auto ready = [&] { return r->state == DONE || r->claimed; };
    if (timeout_ms < 0) {
        s->cv_done.wait(lk, ready);
    } else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
        return fail(MNAD_ERR_TIMEOUT, "ticket %lld still running", (long long)ticket);
    }
    if (r->claimed)
        return fail(MNAD_ERR_STATE, "ticket %lld was already waited for", (long long)ticket);
    r->claimed = true;
    s->all.erase(ticket);
    s->cv_done.notify_all();   // other waiters on this ticket return MNAD_ERR_STATE
    if (info) {
        info->status = r->rc;
        info->submit_s = r->t_submit;
        info->admit_s = r->t_admit;
        info->done_s = r->t_done;
        info->bytes = r->bytes;
    }
    const int rc = r->rc;
    std::string err = r->err;
    if (rc != MNAD_OK) return fail(rc, "ticket %lld: %s", (long long)ticket, err.c_str());
    return MNAD_OK;
}

extern "C" int mnad_cancel(mnad_ctx* ctx, int64_t ticket)
{
    if (!ctx) return fail(MNAD_ERR_ARG, "context is NULL");
    std::shared_ptr<Server> s = get_server(ctx);
    if (!s) return fail(MNAD_ERR_STATE, "the server is not running");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return fail(MNAD_ERR_STATE, "unknown ticket %lld", (long long)ticket);
    Request* r = it->second.get();
    if (r->state == QUEUED) {
        s->queued.erase(std::remove(s->queued.begin(), s->queued.end(), r), s->queued.end());
        r->rc = MNAD_ERR_CANCELLED;
        r->err = "cancelled";
        r->state = DONE;
        r->t_done = now_s();
        s->cv_done.notify_all();
    } else if (r->state == ACTIVE) {
        r->cancel = true;
        s->cv_work.notify_all();
    }
    return MNAD_OK;
}

extern "C" int mnad_serve_status(mnad_ctx* ctx, int* queued, int* active, uint64_t* reserved,
                                 uint64_t* budget, int64_t* steps)
{
    if (!ctx) return fail(MNAD_ERR_ARG, "context is NULL");
    std::shared_ptr<Server> s = get_server(ctx);
    if (!s) return fail(MNAD_ERR_STATE, "the server is not running");
    std::lock_guard<std::mutex> lk(s->mu);
    if (queued) *queued = (int)s->queued.size();
    if (active) {
        int n = 0;
        for (auto& kv : s->all) n += kv.second->state == ACTIVE;
        *active = n;
    }
    if (reserved) *reserved = s->reserved;
    if (budget) *budget = s->budget;
    if (steps) *steps = s->G;
    return MNAD_OK;
}
