// mnad_internal.h -- host-side runtime types shared by the runtime sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "mnad.h"
#include "mnad_device.h"

namespace mnad {

int fail(int code, const char* fmt, ...);
const char* last_error();

// Caching device allocator: freed blocks are kept by exact size and reused
// (repeated shots of one configuration allocate once).
struct Pool {
    std::mutex mu;
    std::multimap<size_t, void*> cache;
    size_t in_use = 0, cached = 0;
    void* get(size_t n, int* rc);
    void put(void* p, size_t n);
    void trim();
};

// Owning handle of one pool block.
struct DevBuf {
    Pool* pool = nullptr;
    void* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : pool(o.pool), p(o.p), n(o.n) { o.p = nullptr; o.n = 0; o.pool = nullptr; }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) {
            reset();
            pool = o.pool; p = o.p; n = o.n;
            o.p = nullptr; o.n = 0; o.pool = nullptr;
        }
        return *this;
    }
    ~DevBuf() { reset(); }
    void reset();
    template <class T> T* as() const { return static_cast<T*>(p); }
};

struct Shot;

// Source wavefield kept between mnad_fwd(indx = 1) and mnad_born_adj: the
// d2u snapshots, or (ring formats) the forward shot itself with its last two
// levels, source amplitudes and saved boundary ring.
struct Kept {
    DevBuf snaps;
    int fmt = 0;
    float sscale = 1.0f;
    int nz = 0, nx = 0, npml = 0, nt = 0, T = 0, K = 0, ncol = 0;
    double dx = 0, dz = 0, dt = 0;
    std::shared_ptr<Shot> fwd;
};

// Autotuner result for one grid shape.
struct TunePick {
    int kernel, rz, brows;   // step-kernel variant
    int batch;               // shots per launch
    double us;               // step-kernel microseconds per shot-step
    // the chosen variant's microseconds per shot-step with 1, 2, 4, ... shots
    // per launch, as far as the search went (0: not measured)
    double curve[8];
    // estimated microseconds per shot-step for n shots per launch (log-linear
    // between measured sizes); a negative value if n lies past the measured ones
    double at(int n) const
    {
        if (n < 1) return -1.0;
        const double x = std::log2((double)n);
        const int k = (int)x;
        if (k >= 8 || curve[k] <= 0) return -1.0;
        if (x == k) return curve[k];
        if (k + 1 >= 8 || curve[k + 1] <= 0) return -1.0;
        return curve[k] + (curve[k + 1] - curve[k]) * (x - k);
    }
};

// One cached launch sequence of run_batch: a graph chunk of step launches
// that reads the batch descriptors from the context's persistent buffers, so
// it can be replayed for any batch with the same kernel and tile count.
// A captured chunk of step launches.  Its key is everything the capture
// bakes into the launches: the kernel and shape, the tile count, the batch's
// shot count, dx == dz (kernel 13) and which passes run around each launch
// (kernel 13), and the device buffers the launches read.
struct GraphEntry {
    int kern, rz, brows, ntiles, chunk, spl, nshots, k1, passes;
    const void* ds;
    const void* dt;
    const int* base;
    hipGraph_t graph;
    hipGraphExec_t exec;
};

// Device arrays that depend on the velocity model only (a / dx^2 per point,
// CPML coefficients), shared by every Model built from the same inputs.
struct ModelBufs {
    DevBuf rx, cfx, cfz;
};

// One entry of the context's model cache: the grid, a 128-bit digest of the
// velocity values and the arrays built from them.
struct ModelKey {
    int nz, nx, npml;
    double dx, dz, dt;
    uint64_t ha, hb;
    bool operator==(const ModelKey& o) const
    {
        return nz == o.nz && nx == o.nx && npml == o.npml && dx == o.dx && dz == o.dz &&
               dt == o.dt && ha == o.ha && hb == o.hb;
    }
};
struct ModelEntry {
    ModelKey key;
    std::shared_ptr<ModelBufs> bufs;
};
constexpr size_t kModelCache = 4;   // models kept per context

// Image / illumination stack (compensated sums over shots).
struct Stack {
    mnad_grid g{};
    DevBuf isum, isumc, lsum, lsumc;
    int nshots = 0;
};

}  // namespace mnad

namespace mnad {
struct Server;
// Default step kernel: wave rows (kernel 4), one wavefront per workgroup.  At
// the C2 size (561 x 1061) one-wavefront workgroups measured 7.57 us per
// one-shot launch against 7.65 / 7.80 / 8.13 us for 2 / 4 / 8 wavefronts.
constexpr int kDefaultKernel = 4, kDefaultRows = 1, kDefaultBrows = 1;
}

struct mnad_ctx {
    int device = 0;
    // serving loop (mnad_serve_start).  Every serving call takes its own
    // reference under srv_mu, so the server and its requests outlive a
    // concurrent mnad_serve_stop until that call returns.
    std::shared_ptr<mnad::Server> server;
    std::mutex srv_mu;
    hipStream_t stream = nullptr;
    std::recursive_mutex mu;
    mnad::Pool pool;
    std::map<int, mnad::Kept> kept;
    std::map<int, mnad::Stack> stacks;
    // most recently used last; a repeated velocity model (every shot of a survey)
    // is validated and hashed on the host but not uploaded or set up again
    std::vector<mnad::ModelEntry> models;
    int next_stack = 1;
    int graph_chunk = 64;
    int kernel = mnad::kDefaultKernel;   // step kernel: 0 registers, 1 prefetch, 2 LDS, 3 two steps,
                                // 4 wave rows, 5 wave rows + one CPML set prefetched, 6 four columns,
                                // 7 two steps, wave rows + LDS
    int rz = mnad::kDefaultRows;       // rows per thread
    int brows = mnad::kDefaultBrows;   // thread rows per block (tile = 64 x brows*rz)
    int strip_first = 1;        // dispatch CPML-strip tiles first
    int ring_order = 3;         // predictor order of the saved boundary ring (2..4)
    int last_group = 0, last_pipelined = 0;   // schedule of the last mnad_rtm_shots
    // XCD balance of one-step kernels: estimated work of a CPML-strip tile and
    // of a four-column tile, in interior one-point tiles
    double strip_weight = 2.0;
    double x4_weight = 2.5;
    size_t budget = 0;          // bytes usable for batches (0 = from free memory)
    int autotune_on_use = 0;    // batched calls tune an unseen grid shape first
    std::map<std::tuple<int, int, int>, mnad::TunePick> tuned;   // (nz, nx, npml)
    // run_batch: persistent batch descriptors (grow-only) and cached graph chunks
    mnad::DevBuf bdesc, btab, bbase;
    mnad::DevBuf bedge;   // resident step kernel: edge records of its blocks
    std::vector<mnad::GraphEntry> graphs;
    // host -> device staging through pinned memory (upload / copy_in): a ring
    // of slots, each reused once the copy that last read it has run
    struct StageSlot {
        void* p = nullptr;
        size_t n = 0;
        hipEvent_t ev = nullptr;
        bool pending = false;
    };
    static constexpr int kStageSlots = 8;
    StageSlot stage[kStageSlots];
    int stage_next = 0;
    // step-kernel accounting (mnad_timing)
    double t_ms = 0.0;
    long long t_launches = 0;
    double t_bytes = 0.0;
};

namespace mnad {

// Velocity-dependent device data of one model.
struct Model {
    mnad_grid g{};
    int pitch = 0, rows = 0;
    size_t elems = 0;              // floats per field (rows * pitch)
    double vmin = 0, vmax = 0;
    const double* vhost = nullptr;  // caller's column-major v, valid during the call
    std::shared_ptr<ModelBufs> bufs;   // shared with the context's model cache
    float* rx() const { return bufs->rx.as<float>(); }
    float* cfx() const { return bufs->cfx.as<float>(); }
    float* cfz() const { return bufs->cfz.as<float>(); }
};

struct ShotSpec {
    int mode = MODE_FWD;
    int nt = 0;
    const double* w = nullptr;   // forward wavelet [nt]
    int zs = -1, xs = -1;
    int nr = 0;
    const int* zr = nullptr;
    const int* xr = nullptr;
    int T = 0, K = 0, ncol = 0;  // snapshots (forward) / imaging (adjoint); T = 0: off
    int snap_fmt = MNAD_SNAP_F32;   // forward: MNAD_SNAP_RING_* saves the boundary ring
    int ring_order = 3;             // predictor order of the ring code (2..4)
};

inline bool ring_format(int fmt) { return fmt == MNAD_SNAP_RING_F16 || fmt == MNAD_SNAP_RING_BF16; }
inline bool snap_format_ok(int fmt) { return fmt >= MNAD_SNAP_F32 && fmt <= MNAD_SNAP_RING_BF16; }

struct Shot {
    ShotDev d{};
    // global step of local step 0: the server's step counter, or the launch
    // of a run_batch sequence at which the shot starts
    long long t0 = 0;
    DevBuf state[12], psx, psz, amp, rrow, rcol, gather, radj, arec, snaps, illum, img, imgc;
    DevBuf rq, rst;                             // saved boundary ring: codes, predictor state
    std::vector<int> uid;                       // receiver r -> unique point u
    std::vector<std::pair<int, int>> upts;      // unique points (j, i)
    std::vector<double> arec_h;                 // (v dt)^2 at unique points
    int nu = 0;
    void release_state()
    {
        for (auto& b : state) b.reset();
        psx.reset();
        psz.reset();
    }
};

int dalloc(mnad_ctx* c, DevBuf& b, size_t bytes, bool zero);
int upload(mnad_ctx* c, DevBuf& b, const void* src, size_t bytes);
// Host -> device copy through the context's pinned staging buffer, queued on
// the context stream; `src` may be reused as soon as this returns.
int copy_in(mnad_ctx* c, void* dst, const void* src, size_t bytes);
// Releases the run_batch descriptors, cached graphs and the staging buffer
// (stream idle).
void release_batch_cache(mnad_ctx* c);
int make_model(mnad_ctx* c, const mnad_grid* g, const double* v, Model& m);
// Host pass over a velocity model: the index of its first value that is not
// positive and finite (n if none) and its range; and the 128-bit digest of
// its values that keys the context's model cache.
size_t scan_model(const double* v, size_t n, double* vmin, double* vmax);
void model_digest(const double* v, size_t n, uint64_t* da, uint64_t* db);
int make_shot(mnad_ctx* c, const Model& m, const ShotSpec& sp, Shot& s);
// Unique receivers of a shot and the kernel's row/column lookup tables (host).
int receiver_tables(const mnad_grid& g, int nr, const int* zr, const int* xr, Shot& s,
                    std::vector<int>& rrow, std::vector<int>& rcol);
// Device address of the allocation-free entry point's scratch array on
// `device` (current), taken once per device; mnad_ctx_create calls it.
int noalloc_init_device(int device);
// Grid checks of make_model (no device work).
int check_grid(const mnad_grid* g);
// Grid part of a shot descriptor (no buffers).
void base_desc(const mnad_grid& g, int pitch, int nt, int mode, int nlev, ShotDev& d);
// Reverse reconstruction of a finished ring-saving forward `fwd` (its last
// two levels, ring, source) as a REC shot paired with the adjoint `adj` of
// the same shot: REC hands d2u to `adj` through a one-snapshot FP32 slot,
// and their launch origins are set so that it does so one launch before the
// adjoint's imaging step reads it.  Takes over fwd's state and ring.
int make_rec_shot(mnad_ctx* c, Shot& fwd, Shot& adj, Shot& rec);
// Kernels that implement the ring code and the reverse reconstruction.
inline bool ring_kernel(int kernel)
{
    return kernel == 4 || kernel == 6 || kernel == 10 || kernel == 11 || kernel == 13;
}
void attach_snaps(Shot& adj, DevBuf&& snaps, int fmt, float sscale);
int fill_radj_host(mnad_ctx* c, Shot& s, int nt, int nr, const double* data);
int run_batch(mnad_ctx* c, const std::vector<Shot*>& shots, int flags);
// Step kernel 13: estimated cost of an edge wave (bounds checks, CPML
// memories of both sets) against an interior wave of the same size, from
// kernel traces of separate interior and edge launches at the 2061 x 8061
// grid (2.35x per point); build_tiles balances the XCDs' work with it.
constexpr double kQEdgeWeight = 2.4;
void build_tiles(const mnad_ctx* c, std::vector<ShotDev>& h, std::vector<TileRef>& table);
// Step kernels that advance two time levels per launch (shots keep four).
inline bool two_step_kernel(int kernel)
{
    return kernel == 7 || kernel == 13;
}
// Columns per lane of the interior tiles of kernels 6 (4) and 10 (2); 0 for others.
int interior_columns_per_lane(int kernel);
// Step kernel that runs a whole batch in one launch (resident blocks).
inline bool resident_kernel(int kernel) { return kernel == 9; }

// Output tile of the current step kernel (columns x rows) and steps per launch.
void tile_shape(const mnad_ctx* c, int* tw, int* th, int* steps_per_launch);
double algorithmic_bytes(const ShotDev& d);
// Downloads run on stream st (default: the context's stream) and wait for it.
int download_gather(mnad_ctx* c, const Shot& s, int nt, int nr, double* out,
                    hipStream_t st = nullptr);
int download_field(mnad_ctx* c, const Model& m, const Shot& s, int field, double scale,
                   double* out);
int download_interior(mnad_ctx* c, const mnad_grid& g, const float* a, const float* b,
                      double scale, double* out, hipStream_t st = nullptr);
int schedule(int nt, double vmin, double vmax, int* nf, int* T, int* K, int* ncol);
int autotune(mnad_ctx* c, const Model& m, int max_shots, TunePick* out);
void server_destroy(mnad_ctx* c);
const TunePick* find_tuned(mnad_ctx* c, const mnad_grid& g);

// Applies a tuned step-kernel variant for the lifetime of the guard.
struct TuningGuard {
    mnad_ctx* c;
    int k, r, b;
    TuningGuard(mnad_ctx* c_, const TunePick* p) : c(c_), k(c_->kernel), r(c_->rz), b(c_->brows)
    {
        if (p) {
            c->kernel = p->kernel;
            c->rz = p->rz;
            c->brows = p->brows;
        }
    }
    ~TuningGuard()
    {
        c->kernel = k;
        c->rz = r;
        c->brows = b;
    }
};

}  // namespace mnad
