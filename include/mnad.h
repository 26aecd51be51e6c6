/*
 * mnad.h -- C-ABI of libmnad, the MI355X (gfx950) implementation of the
 * MNAD-MRTM hot path: the coupled (u, du/dx, du/dz) MNAD time step with CPML,
 * source injection, receiver sampling, RTM receiver back-propagation and the
 * zero-lag imaging condition.
 *
 * Plain C types only.  Array conventions follow what a MATLAB MEX gateway
 * hands over (mxGetPr): column-major, z (row) index fastest.
 *   v        [nz x nx]   padded velocity, v(j,i) = v[j + i*nz]
 *   w        [nt]        source wavelet
 *   datau    [nt x nr]   gather, datau(t,r) = datau[t + r*nt]
 *   U, V, W  [nz x nx]   final fields (V = dU/dx-like field, W = z analogue,
 *                        exactly as fwd_PDE_exact_MNAD.m carries them)
 *   illum    [nzi x nxi] interior illumination, nzi = nz - 2 npml
 *   Im       [nzi * nxi] interior image, column-major (Born_adj_MNAD.m:138)
 * Grid indices (zs, xs, zr, xr) are 0-based padded indices; the MATLAB
 * reference passes 1-based ones (main_Acoustic_Data.m:120-126), so a gateway
 * subtracts 1.
 *
 * Every function returns MNAD_OK (0) or a negative MNAD_ERR_* code; the text
 * of the last error on the calling thread is returned by mnad_last_error().
 * All entry points are re-entrant: calls on one context are serialised.
 */
#ifndef MNAD_H
#define MNAD_H

#include <stdint.h>

#if defined(__GNUC__)
#define MNAD_API __attribute__((visibility("default")))
#else
#define MNAD_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define MNAD_OK 0
#define MNAD_ERR_ARG (-1)       /* invalid argument (shape, index, CFL, NaN) */
#define MNAD_ERR_HIP (-2)       /* HIP runtime / device error */
#define MNAD_ERR_NOMEM (-3)     /* device memory exhausted */
#define MNAD_ERR_STATE (-4)     /* missing source wavefield, unknown handle */
#define MNAD_ERR_CANCELLED (-5) /* request cancelled */
#define MNAD_ERR_BUDGET (-6)    /* request larger than the HBM budget */
#define MNAD_ERR_NOGPU (-7)     /* no HIP device */
#define MNAD_ERR_TIMEOUT (-8)   /* wait timed out (request still running) */

/* Storage of the source wavefield an RTM forward keeps for its adjoint.
 * F32 / F16 / BF16: the d2u snapshots of fwd_PDE_exact_MNAD.m:126-143 every T
 * steps, as the reference's UU (K x interior points).
 * RING_F16 / RING_BF16: the forward's last two levels plus (U, dU/dx, dU/dz)
 * on a ring of two cells around the interior every step, coded as 16-bit
 * residuals of a closed-loop predictor; the adjoint recomputes the source
 * wavefield backwards inside the CPML and takes d2u from it.  About
 * 12 (nx + nz) bytes per step instead of 2 nx nz / T. */
#define MNAD_SNAP_F32 0
#define MNAD_SNAP_F16 1   /* FP16 with a per-shot power-of-two scale */
#define MNAD_SNAP_BF16 2
#define MNAD_SNAP_RING_F16 3
#define MNAD_SNAP_RING_BF16 4

/* Flags */
#define MNAD_FLAG_NO_GRAPH 1   /* plain launches, one step per launch */
#define MNAD_FLAG_SERIAL 2     /* one shot and one propagation per launch sequence */
#define MNAD_FLAG_AUTOTUNE 4   /* batched calls: autotune an unseen grid shape first */

typedef struct {
    int nz, nx;       /* padded grid (rows = z, columns = x) */
    int npml;         /* CPML thickness in cells (d_pml) */
    double dx, dz, dt;
} mnad_grid;

typedef struct mnad_ctx mnad_ctx;

/* Library identification and last error of the calling thread. */
MNAD_API const char *mnad_version(void);
MNAD_API const char *mnad_last_error(void);

/* Number of visible HIP devices (0 when none). */
MNAD_API int mnad_device_count(void);

/* Create / destroy a context bound to one device.  A context owns the
 * device stream, the memory pool and the source-wavefield handles. */
MNAD_API int mnad_ctx_create(int device, mnad_ctx **out);
MNAD_API void mnad_ctx_destroy(mnad_ctx *ctx);
/* Waits until all work queued on the context's stream has finished (every
 * other entry point returns with its results on the host already). */
MNAD_API int mnad_ctx_sync(mnad_ctx *ctx);

/* Forward shot.  Replaces
 *   data = fwd_PDE_exact_MNAD(dx,xs,xr,dz,zs,zr,v,dt,nt,w,d_pml,ss,Auu,...,Aww,indx)
 * (SRC/MNAD/fwd_PDE_exact_MNAD.m:1).  The nine sparse operators are not
 * passed: the kernel evaluates their entries from v, dx, dz, dt.
 * indx = 1 keeps the source-side d2u snapshots on the device under the key
 * ss (replacing ./Wavefields/UU_<ss>.mat, fwd_PDE_exact_MNAD.m:161) and
 * returns T*sum(d2u^2) in `illum` when non-NULL (./Illuminations/illum_<ss>,
 * fwd_PDE_exact_MNAD.m:162-163).  U, V, W, illum may be NULL. */
MNAD_API int mnad_fwd(mnad_ctx *ctx, const mnad_grid *g, const double *v, int nt,
             const double *w, int zs, int xs, int nr, const int *zr,
             const int *xr, int ss, int indx, int snap_fmt, int flags,
             double *datau, double *U, double *V, double *W, double *illum);

/* RTM adjoint of one shot.  Replaces
 *   Im = Born_adj_MNAD(dx,dz,nt,dt,v,xr,zr,data,d_pml,ss,Auu,...,Aww)
 * (SRC/migration/Born_adj_MNAD.m:1) with U_r read as U and `data` a matrix
 * [nt x nr].  Consumes the snapshots kept by mnad_fwd(..., ss, indx=1) on the
 * same v (MNAD_ERR_STATE if there are none) and releases them. */
MNAD_API int mnad_born_adj(mnad_ctx *ctx, const mnad_grid *g, const double *v, int nt,
                  int nr, const int *zr, const int *xr, const double *data,
                  int ss, int flags, double *Im);

/* Drop the snapshots kept for shot ss (no-op if none). */
MNAD_API int mnad_release(mnad_ctx *ctx, int ss);

/* Spill the snapshots kept for shot ss to a file / load them back as the
 * kept snapshots of ss.  This is the cross-process hand-off the reference
 * does with ./Wavefields/UU_<ss>.mat (fwd_PDE_exact_MNAD.m:161,
 * Born_adj_MNAD.m:33) when the forward and the adjoint of one shot run in
 * different parfor workers.  The file holds a small header (grid, nt, T, K,
 * ncol, format, scale) and the snapshots in their storage format. */
MNAD_API int mnad_snap_save(mnad_ctx *ctx, int ss, const char *path);
MNAD_API int mnad_snap_load(mnad_ctx *ctx, int ss, const char *path);

/* Snapshot schedule of fwd_PDE_exact_MNAD.m:46-47 for a velocity range. */
MNAD_API int mnad_snap_schedule(int nt, double vmin, double vmax, int *nf, int *T,
                       int *K, int *ncol);

/* ---- RTM workflow on the device (main_Acoustic_RTM.m:172-262) ----------
 * A stack accumulates, over shots, the per-shot images Im_s and
 * illuminations illum_s with compensated FP32 sums; mnad_stack_image returns
 * img = mean_s(Im_s) ./ mean_s(illum_s) (main_Acoustic_RTM.m:243,262). */
MNAD_API int mnad_stack_create(mnad_ctx *ctx, const mnad_grid *g, int *stack_id);
MNAD_API int mnad_stack_destroy(mnad_ctx *ctx, int stack_id);
MNAD_API int mnad_stack_image(mnad_ctx *ctx, int stack_id, double *img, double *illum,
                     int *nshots);

/* One receiver spread per shot: zs[k], xs[k] and rec_off[k]..rec_off[k+1]-1
 * into zr/xr (rec_off has nshots+1 entries). */
typedef struct {
    int nshots;
    const int *zs, *xs;
    const int *rec_off;
    const int *zr, *xr;
} mnad_shots;

/* RTM of several shots, batched: per shot, a forward on v_true (observed
 * gather), a forward on v_smooth with snapshots (synthetic gather), the
 * scattered data d_obs - d_syn (main_Acoustic_RTM.m:225) and the adjoint with
 * imaging (Born_adj_MNAD.m); images and illuminations go into the stack.
 * Up to `batch` shots share one launch sequence (0 = choose from the memory
 * budget).  dobs / dsyn (may be NULL) receive the gathers of every shot,
 * concatenated per shot, each [nt x nr_k] column-major. */
MNAD_API int mnad_rtm_shots(mnad_ctx *ctx, int stack_id, const mnad_grid *g,
                   const double *v_true, const double *v_smooth, int nt,
                   const double *w, const mnad_shots *shots, int snap_fmt,
                   int batch, int flags, double *dobs, double *dsyn);

/* Schedule the last mnad_rtm_shots call on this context used: shots per
 * group (one launch sequence runs a group's forwards, the next its adjoints)
 * and whether a group's forwards shared launches with the previous group's
 * adjoints (pipelined = 1) or ran in order (0). */
MNAD_API int mnad_rtm_schedule(mnad_ctx *ctx, int *shots_per_group, int *pipelined);

/* Forward modelling of several shots on one model, batched in one launch
 * sequence (the parfor of main_Acoustic_Data.m:169-171).  datau receives the
 * gathers concatenated per shot ([nt x nr_k] column-major each). */
MNAD_API int mnad_fwd_shots(mnad_ctx *ctx, const mnad_grid *g, const double *v, int nt,
                   const double *w, const mnad_shots *shots, int batch,
                   int flags, double *datau);

/* Forward shot without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The same computation as mnad_fwd(..., indx = 0, flags = MNAD_FLAG_NO_GRAPH)
 * and the same results; every device buffer beyond the caller's inputs and
 * outputs comes from a fixed 64 MiB array in device memory that each GPU the
 * library runs on holds until the process ends (about a single 1001 x 501
 * shot with 30 CPML cells, 4000 steps and 1001 receivers; MNAD_ERR_ARG, with
 * the bytes needed, beyond that).  Calls on one GPU share that array: within
 * a process they are serialised, and each call clears what must start at
 * zero.  Inputs and outputs are host arrays as in mnad_fwd, copied on the
 * context's stream; U, V, W, datau may be NULL. */
MNAD_API int mnad_fwd_noalloc(mnad_ctx *ctx, const mnad_grid *g, const double *v, int nt,
                              const double *w, int zs, int xs, int nr, const int *zr,
                              const int *xr, double *datau, double *U, double *V, double *W);

/* Step-kernel accounting since the last reset: GPU time of the step-kernel
 * launch sequences measured with HIP events on the context's stream (ms,
 * inter-launch gaps included), the number of step launches, and the
 * algorithmic HBM bytes those launches move (40 B per grid point, 84 B per
 * CPML strip point and direction: nine memories read and written, three
 * damping coefficients read; receivers, snapshots, image).  reset != 0
 * clears the counters after reading. */
MNAD_API int mnad_timing(mnad_ctx *ctx, int reset, double *ms, int64_t *launches,
                         double *bytes);

/* Launch-shape knobs of the step kernel (chosen by the autotuner):
 * a tile is 64 x (block_rows * rows_per_thread) points, one block per tile;
 * kernel 0 = neighbourhoods loaded through the vector cache: block_rows 4
 *   with rows_per_thread 1, 2 or 4, or block_rows 8 or 16 with 1 row;
 * kernel 4 (default) = one grid row per wavefront, row origins in scalar
 *   registers and one shared lane offset for every load and store: block_rows
 *   1 (default), 2, 4 or 8, 1 row;
 * kernel 6 = kernel 4 on the CPML strips, four columns per lane (16-byte
 *   loads, neighbours by cross-lane shifts) in the interior: block_rows 1, 2
 *   or 4, 1 row;
 * kernel 7 = two time steps per launch, level s staged in LDS with a 4-point
 *   halo, one grid row per wavefront: 60 x (4 rows_per_thread) output tiles
 *   with rows_per_thread 3, 5 or 7 and block_rows (wavefronts) 4, 8 or 16
 *   dividing 4 (rows_per_thread + 1); shots keep four time levels -- set it
 *   before creating shots;
 * kernel 10 = kernel 6 with two columns per lane in the interior (8-byte
 *   loads, 124 columns per wave, fewer registers): block_rows 1, 2 or 4, 1 row;
 * kernel 11 = two grid rows per wavefront (the rows both updates need are
 *   loaded once): block_rows 1, rows_per_thread 2;
 * kernel 13 = two time steps per launch, each wavefront marching 56 columns
 *   down a segment of 8 rows_per_thread rows (1-256) with the rows it needs
 *   in register rings; block_rows (1, 2 or 4) wavefronts side by side per
 *   workgroup; shots keep four time levels -- set it before creating shots;
 *   for large batches (a launch waits for one wavefront's whole march);
 * kernel 9 = resident blocks: one cooperative launch runs every step of a
 *   batch; a workgroup of 8 wavefronts keeps a 64-column block of one shot in
 *   LDS for the whole launch and exchanges only its two-cell edges with its
 *   neighbours (agent-scope atomic records, no per-step launch or cache
 *   flush).  block_rows 8; rows_per_thread = blocks per CU wanted (1-4; fewer
 *   are used when the batch does not fit, MNAD_ERR_ARG when no block height up
 *   to 64 rows fits);
 * graph_chunk = step launches per captured hipGraph (<= 1: no graph);
 * strip_first != 0 dispatches the CPML-strip tiles of each XCD first. */
typedef struct {
    int kernel;
    int rows_per_thread;
    int block_rows;
    int graph_chunk;
    int strip_first;
} mnad_tuning;
MNAD_API int mnad_get_tuning(mnad_ctx *ctx, mnad_tuning *t);
MNAD_API int mnad_set_tuning(mnad_ctx *ctx, const mnad_tuning *t);

/* ---- Serving: continuous batching of requests of any grid size ---------
 * A server thread owns a private stream.  Requests are copied at submit and
 * queued; at every chunk boundary (chunk_steps time steps) the loop admits
 * queued requests in submission order while the HBM budget, max_active and
 * max_points allow, runs every active shot of every
 * admitted request in one launch per step, and completes requests whose
 * last step has run.  An RTM request runs its two forward propagations
 * (v, v_smooth with snapshots), then its adjoint with imaging, in place.
 * Results do not depend on what else is in the batch.  Output buffers must
 * stay valid until mnad_wait returns for the ticket. */
#define MNAD_REQ_MODEL 0   /* forward modelling on v: out_gather */
#define MNAD_REQ_RTM 1     /* modelling on v + RTM on v_smooth: out_image */

typedef struct {
    int kind;                  /* MNAD_REQ_MODEL or MNAD_REQ_RTM */
    mnad_grid grid;
    const double *v;           /* [nz x nx] column-major */
    const double *v_smooth;    /* RTM migration model, NULL for modelling */
    int nt;
    const double *w;           /* [nt] */
    int zs, xs;                /* 0-based padded source */
    int nr;
    const int *zr, *xr;        /* 0-based padded receivers */
    int snap_fmt;              /* RTM snapshot storage */
    double *out_gather;        /* [nt x nr] observed gather (optional for RTM) */
    double *out_image;         /* RTM: Im (Born_adj_MNAD.m:138), [nzi x nxi] */
    double *out_illum;         /* RTM: T*sum(d2u^2), [nzi x nxi] (optional) */
} mnad_request;

typedef struct {
    int status;                /* MNAD_OK or the request's error code */
    double submit_s, admit_s, done_s;   /* server clock, seconds */
    uint64_t bytes;            /* device bytes reserved for the request */
} mnad_req_info;

/* hbm_budget: device bytes the admitted requests may reserve (0: 80 % of free
 * memory); max_active: requests in flight at most; max_points: grid points
 * (summed over the shots stepping) above which no further request is
 * admitted -- enough to fill the GPU without stretching every request by
 * sharing it more widely (0: no limit); chunk_steps: steps between
 * admissions.  Admission is first come, first served. */
MNAD_API int mnad_serve_start(mnad_ctx *ctx, uint64_t hbm_budget, int max_active,
                              uint64_t max_points, int chunk_steps);
MNAD_API int mnad_serve_stop(mnad_ctx *ctx);
MNAD_API int mnad_submit(mnad_ctx *ctx, const mnad_request *req, int64_t *ticket);
/* timeout_ms < 0 waits forever; MNAD_ERR_TIMEOUT leaves the ticket pending. */
MNAD_API int mnad_wait(mnad_ctx *ctx, int64_t ticket, int timeout_ms, mnad_req_info *info);
MNAD_API int mnad_cancel(mnad_ctx *ctx, int64_t ticket);
/* queued / active requests, reserved and budget bytes, steps run so far */
MNAD_API int mnad_serve_status(mnad_ctx *ctx, int *queued, int *active, uint64_t *reserved,
                               uint64_t *budget, int64_t *steps);

/* Autotuner: times the step-kernel variants and 1, 2, 4, ... up to max_shots
 * shots per launch with short forward probes on this grid and velocity, keeps
 * the fastest per (nz, nx, npml) in the context and returns it.  Batched calls
 * with batch = 0 then use the tuned variant and shots per launch (an RTM
 * group of B shots puts 2B forward shots in one launch). */
MNAD_API int mnad_autotune(mnad_ctx *ctx, const mnad_grid *g, const double *v,
                           int max_shots, mnad_tuning *best, int *best_shots,
                           double *us_per_shot_step);

/* Device memory in use / cached by the context's pool, in bytes. */
MNAD_API int mnad_mem_info(mnad_ctx *ctx, uint64_t *in_use, uint64_t *cached);

#ifdef __cplusplus
}
#endif
#endif
