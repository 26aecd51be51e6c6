// mnad_autotune.hip -- picks the step-kernel variant and the number of shots
// per launch for one grid shape by timing short forward probes on the device.
//
// Why measure: on MI355X the best batch depends on whether a launch's
// working set (about 40 B per point per shot and step) fits the 256 MB
// Infinity Cache; a single 3401 x 751 shot (130 MB per step) already runs
// from it, while 595 k-point shots need several per launch to fill the
// 256 CUs.  The kernel variant trades parallelism against memory
// instructions the same way: one grid row per wavefront (most wavefronts,
// best latency hiding), four columns per lane (a quarter of the load
// instructions and redundant row reads, half the occupancy), two steps per
// launch (half the launches and level traffic, redundant halo arithmetic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "mnad.h"
#include "mnad_internal.h"

namespace mnad {

namespace {

struct Variant {
    int kernel, rz, brows;
};

const Variant kVariants[] = {{4, 1, 4}, {4, 1, 1}, {11, 2, 1}, {0, 1, 4}, {6, 1, 2}, {6, 1, 4},
                             {10, 1, 2}, {7, 3, 16}, {7, 5, 8}, {7, 7, 16}, {13, 2, 4},
                             {13, 4, 4}};

// Snapshot period of an RTM forward on this model: T = floor(nt / nf / 2) with
// nf = floor(nt vmin / (6 vmax)), about 3 vmax / vmin.
int probe_period(const Model& m) { return std::max(3, (int)(3.0 * m.vmax / m.vmin)); }

// Time `steps` steps of `b` forward shots (source at the centre, no
// receivers) with the context's current tuning; returns microseconds per
// shot-step of step-kernel time.  Like an RTM launch, about two shots in
// three take f16 d2u snapshots every probe_period steps (the adjoint's
// imaging term has the same shape), so that a variant whose snapshot or
// imaging steps cost more is seen to.
int probe(mnad_ctx* c, const Model& m, int b, int steps, double* us)
{
    std::vector<double> w(steps, 0.0);
    for (int t = 0; t < steps; ++t) w[t] = std::sin(0.3 * t) * std::exp(-0.01 * t);
    std::vector<Shot> shots(b);
    std::vector<Shot*> run;
    const int T = probe_period(m);
    for (int k = 0; k < b; ++k) {
        ShotSpec sp;
        sp.mode = MODE_FWD;
        sp.nt = steps;
        sp.w = w.data();
        sp.zs = m.g.nz / 2;
        sp.xs = m.g.nx / 2;
        if (k % 3 != 2 && steps >= T) {
            sp.T = T;
            sp.K = steps / T;
            sp.ncol = sp.K;
            sp.snap_fmt = MNAD_SNAP_F16;
        }
        int rc = make_shot(c, m, sp, shots[k]);
        if (rc != MNAD_OK) return rc;
        run.push_back(&shots[k]);
    }
    const double ms0 = c->t_ms;
    const long long n0 = c->t_launches;
    const double b0 = c->t_bytes;
    int rc = run_batch(c, run, 0);
    // probes do not count as work in the context's accounting
    const double ms = c->t_ms - ms0;
    c->t_ms = ms0;
    c->t_launches = n0;
    c->t_bytes = b0;
    if (rc != MNAD_OK) return rc;
    *us = ms * 1e3 / ((double)b * steps);
    return MNAD_OK;
}

}  // namespace

int autotune(mnad_ctx* c, const Model& m, int max_shots, TunePick* out)
{
    const mnad_grid& g = m.g;
    // probe length: enough launches to amortise the graph chunking, fewer on big grids
    const double pts = (double)g.nz * g.nx;
    const int steps = pts > 4e6 ? 2 * c->graph_chunk : 4 * c->graph_chunk;
    // shots per launch: powers of two up to max_shots, bounded by memory.  A
    // probe shot holds 3 nlev field planes and 9 (one-step kernels) or 18
    // (two-step kernels: two sets) CPML memory planes of its strips, plus its
    // snapshots and illumination.
    const size_t ni = (size_t)std::max(0, g.nz - 2 * g.npml) * std::max(0, g.nx - 2 * g.npml);
    const size_t strip = (size_t)g.nz * 2 * g.npml + (size_t)2 * g.npml * m.pitch;
    auto per_shot = [&](int kernel) {
        const size_t nlev = two_step_kernel(kernel) ? 4 : 2;
        return 3 * nlev * m.elems * 4 + 9 * (nlev / 2) * strip * 4 +
               (size_t)(steps / probe_period(m)) * ni * 2 + ni * 4 + 4096;
    };
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) {
        (void)hipGetLastError();
        freeb = 0;
    }
    const double avail = 0.5 * (double)(freeb + c->pool.cached);
    auto mem_cap = [&](int kernel) {
        return (int)std::max<size_t>(1, (size_t)(avail / (double)per_shot(kernel)));
    };
    max_shots = std::max(1, std::min(max_shots, mem_cap(4)));

    const int sk = c->kernel, sr = c->rz, sb = c->brows;
    const bool log = getenv("MNAD_AUTOTUNE_LOG") != nullptr;   // one stderr line per probe
    // MNAD_AUTOTUNE_SKIP: comma-separated kernel numbers left out of the search
    std::vector<int> skip;
    if (const char* e = getenv("MNAD_AUTOTUNE_SKIP")) {
        for (const char* q = e; *q;) {
            char* end = nullptr;
            const long k = strtol(q, &end, 10);
            if (end == q) break;
            skip.push_back((int)k);
            q = (*end == ',') ? end + 1 : end;
        }
    }
    TunePick best{0, 1, 4, 1, 1e30, {}};
    constexpr int nv = (int)(sizeof(kVariants) / sizeof(kVariants[0]));
    double curves[nv][8] = {};   // per variant: us per shot-step at 1, 2, 4, ... shots
    int kbest = 0;
    int ran = 0, first_rc = MNAD_OK;
    std::string first_err;
    for (int b = 1, kb = 0; b <= max_shots && kb < 8; b *= 2, ++kb) {
        double best_b = 1e30;
        for (int iv = 0; iv < nv; ++iv) {
            const Variant& v = kVariants[iv];
            curves[iv][kb] = 1e30;
            if (std::find(skip.begin(), skip.end(), v.kernel) != skip.end()) continue;
            if (b > mem_cap(v.kernel)) continue;   // this variant's shots do not fit
            c->kernel = v.kernel;
            c->rz = v.rz;
            c->brows = v.brows;
            double us = 0;
            int rc = probe(c, m, b, steps, &us);   // warm-up: first touch, code load
            if (rc == MNAD_OK) rc = probe(c, m, b, steps, &us);
            if (rc != MNAD_OK) {
                // a variant that cannot run here (memory, launch limits) is left out
                if (first_rc == MNAD_OK) {
                    first_rc = rc;
                    first_err = last_error();
                }
                if (log)
                    fprintf(stderr, "mnad_autotune %dx%d: kernel %d rows %d block_rows %d, %d shots: "
                            "skipped (%s)\n", g.nz, g.nx, v.kernel, v.rz, v.brows, b, last_error());
                continue;
            }
            ++ran;
            if (log)
                fprintf(stderr, "mnad_autotune %dx%d: kernel %d rows %d block_rows %d, %d shots: %.3f us/shot-step\n",
                        g.nz, g.nx, v.kernel, v.rz, v.brows, b, us);
            best_b = std::min(best_b, us);
            curves[iv][kb] = us;
            if (us < best.us) {
                best = TunePick{v.kernel, v.rz, v.brows, b, us, {}};
                kbest = iv;
            }
        }
        // stop growing the batch once it no longer helps (working set past the cache)
        if (b > 1 && best_b > 1.05 * best.us) break;
    }
    c->kernel = sk;
    c->rz = sr;
    c->brows = sb;
    if (ran == 0)
        return fail(first_rc == MNAD_OK ? MNAD_ERR_ARG : first_rc,
                    "autotune %dx%d: no step-kernel variant ran (%s)", g.nz, g.nx,
                    first_rc == MNAD_OK ? "all skipped" : first_err.c_str());
    for (int k = 0; k < 8; ++k) best.curve[k] = curves[kbest][k] < 1e29 ? curves[kbest][k] : 0.0;
    *out = best;
    c->tuned[std::make_tuple(g.nz, g.nx, g.npml)] = best;
    return MNAD_OK;
}

const TunePick* find_tuned(mnad_ctx* c, const mnad_grid& g)
{
    auto it = c->tuned.find(std::make_tuple(g.nz, g.nx, g.npml));
    return it == c->tuned.end() ? nullptr : &it->second;
}

}  // namespace mnad

extern "C" int mnad_autotune(mnad_ctx* ctx, const mnad_grid* g, const double* v, int max_shots,
                             mnad_tuning* best, int* best_shots, double* us_per_shot_step)
{
    using namespace mnad;
    if (!ctx) return fail(MNAD_ERR_ARG, "context is NULL");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(MNAD_ERR_HIP, "hipSetDevice failed");
    Model m;
    int rc = make_model(ctx, g, v, m);
    if (rc != MNAD_OK) return rc;
    TunePick p;
    rc = autotune(ctx, m, std::max(1, max_shots), &p);
    if (rc != MNAD_OK) return rc;
    if (best) {
        best->kernel = p.kernel;
        best->rows_per_thread = p.rz;
        best->block_rows = p.brows;
        best->graph_chunk = ctx->graph_chunk;
        best->strip_first = ctx->strip_first;
    }
    if (best_shots) *best_shots = p.batch;
    if (us_per_shot_step) *us_per_shot_step = p.us;
    return MNAD_OK;
}
