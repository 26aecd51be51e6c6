/*
 * mnad_mex.c -- MATLAB MEX gateway over libmnad (include/mnad.h).
 *
 * Build inside MATLAB, with libmnad.so on the loader path at run time:
 *   mex -R2017b -I<repo>/include mnad_mex.c -L<repo>/mnad-modeling-rtm_amd/_build -lmnad
 *
 * Commands (first argument):
 *   datau          = mnad_mex('fwd', dx,dz,dt,d_pml, v, w, nt, xs,zs, xr,zr, ss, indx)
 *   [datau, illum] = mnad_mex('fwd', ...)           (illum only with indx == 1)
 *   Im             = mnad_mex('adj', dx,dz,nt,dt, v, xr,zr, data, d_pml, ss)
 *   mnad_mex('save', ss, path) / mnad_mex('load', ss, path) / mnad_mex('release', ss)
 *   s              = mnad_mex('version')
 * Many shots in one call (the parfor loops of main_Acoustic_Data.m:169-171 and
 * main_Acoustic_RTM.m:172-262 as the batch axis); shot k has nrec(k)
 * receivers, xr/zr hold every shot's receivers one after the other:
 *   datau = mnad_mex('fwd_shots', dx,dz,dt,d_pml, v, w, nt, xs,zs, xr,zr, nrec, batch)
 *   [img, illum, dobs, dsyn] = mnad_mex('rtm_shots', dx,dz,dt,d_pml, v, vs, w, nt,
 *                                       xs,zs, xr,zr, nrec, batch)
 *     img = mean(Im_s) ./ mean(illum_s) [nzi x nxi] (main_Acoustic_RTM.m:262);
 *     datau, dobs, dsyn: [nt x sum(nrec)], the shots' gathers side by side.
 * Serving (continuous batching of requests of any grid size):
 *   mnad_mex('serve_start', budget_bytes, max_active, max_points, chunk_steps)
 *   t = mnad_mex('submit', kind, dx,dz,dt,d_pml, v, vs, w, nt, xs,zs, xr,zr)
 *         kind 0: modelling on v (vs = []), 1: RTM on v and vs
 *   [out, status, times] = mnad_mex('wait', t, timeout_ms)
 *         out: the gather [nt x nr] (modelling) or Im [nzi x nxi] (RTM);
 *         times = [submit, admit, done] seconds on the server clock
 *   mnad_mex('cancel', t) / mnad_mex('serve_stop')
 * The fwd/adj argument lists are those of fwd_PDE_exact_MNAD.m:1 and
 * Born_adj_MNAD.m:1 without the nine operator matrices; grid indices are the
 * reference's 1-based padded indices (main_Acoustic_Data.m:120-126).
 * Source-wavefield storage of RTM: f16 snapshots, or MNAD_SNAP=f32|bf16|
 * ring_f16|ring_bf16 in the environment (ring_*: the boundary ring with
 * reverse reconstruction, about 15 times less device memory per shot).
 * Errors raise mnad:<kind> identifiers (mnad:arg, mnad:hip, mnad:nomem,
 * mnad:state, mnad:nogpu, ...).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"
#include "mnad.h"

static mnad_ctx *g_ctx = NULL;

static void cleanup(void)
{
    if (g_ctx) mnad_ctx_destroy(g_ctx);   /* stops a running server first */
    g_ctx = NULL;
}

static const char *err_id(int rc)
{
    switch (rc) {
    case MNAD_ERR_ARG: return "mnad:arg";
    case MNAD_ERR_HIP: return "mnad:hip";
    case MNAD_ERR_NOMEM: return "mnad:nomem";
    case MNAD_ERR_STATE: return "mnad:state";
    case MNAD_ERR_CANCELLED: return "mnad:cancelled";
    case MNAD_ERR_BUDGET: return "mnad:budget";
    case MNAD_ERR_NOGPU: return "mnad:nogpu";
    case MNAD_ERR_TIMEOUT: return "mnad:timeout";
    default: return "mnad:error";
    }
}

static void check(int rc)
{
    if (rc != MNAD_OK) mexErrMsgIdAndTxt(err_id(rc), "%s", mnad_last_error());
}

static mnad_ctx *context(void)
{
    if (!g_ctx) {
        check(mnad_ctx_create(0, &g_ctx));
        mexAtExit(cleanup);
    }
    return g_ctx;
}

static int snap_format(void)
{
    const char *s = getenv("MNAD_SNAP");
    if (s && strcmp(s, "f32") == 0) return MNAD_SNAP_F32;
    if (s && strcmp(s, "bf16") == 0) return MNAD_SNAP_BF16;
    if (s && strcmp(s, "ring_f16") == 0) return MNAD_SNAP_RING_F16;
    if (s && strcmp(s, "ring_bf16") == 0) return MNAD_SNAP_RING_BF16;
    return MNAD_SNAP_F16;
}

static const double *real_double(const mxArray *a, const char *name)
{
    if (!mxIsDouble(a) || mxIsComplex(a) || mxIsSparse(a))
        mexErrMsgIdAndTxt("mnad:arg", "%s must be a real full double array", name);
    return mxGetPr(a);
}

static double scalar(const mxArray *a, const char *name)
{
    if (mxGetNumberOfElements(a) != 1) mexErrMsgIdAndTxt("mnad:arg", "%s must be a scalar", name);
    return real_double(a, name)[0];
}

static int index0(double x, const char *name)
{
    if (!(x == floor(x)) || x < 1 || x > 2147483647.0)
        mexErrMsgIdAndTxt("mnad:arg", "%s must hold positive integer grid indices", name);
    return (int)x - 1;
}

static int as_int(double x, const char *name)
{
    if (!(x == floor(x)) || fabs(x) > 2147483647.0)
        mexErrMsgIdAndTxt("mnad:arg", "%s must be an integer", name);
    return (int)x;
}

/* 1-based MATLAB index vector -> 0-based int array (mxMalloc'd). */
static int *indices(const mxArray *a, const char *name, size_t *n)
{
    const double *p = real_double(a, name);
    size_t k, m = mxGetNumberOfElements(a);
    int *out = (int *)mxMalloc((m ? m : 1) * sizeof(int));
    for (k = 0; k < m; ++k) out[k] = index0(p[k], name);
    *n = m;
    return out;
}

static void cmd_fwd(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    mnad_grid g;
    const double *v, *w;
    int nt, xs, zs, ss, indx;
    size_t nr, nr2;
    int *xr, *zr;
    double *illum = NULL;
    if (nrhs != 14)
        mexErrMsgIdAndTxt("mnad:arg", "usage: mnad_mex('fwd', dx,dz,dt,d_pml, v, w, nt, xs,zs, xr,zr, ss, indx)");
    g.dx = scalar(prhs[1], "dx");
    g.dz = scalar(prhs[2], "dz");
    g.dt = scalar(prhs[3], "dt");
    g.npml = as_int(scalar(prhs[4], "d_pml"), "d_pml");
    v = real_double(prhs[5], "v");
    g.nz = (int)mxGetM(prhs[5]);
    g.nx = (int)mxGetN(prhs[5]);
    w = real_double(prhs[6], "w");
    nt = as_int(scalar(prhs[7], "nt"), "nt");
    if ((size_t)nt != mxGetNumberOfElements(prhs[6]))
        mexErrMsgIdAndTxt("mnad:arg", "length(w) = %d differs from nt = %d",
                          (int)mxGetNumberOfElements(prhs[6]), nt);
    xs = index0(scalar(prhs[8], "xs"), "xs");
    zs = index0(scalar(prhs[9], "zs"), "zs");
    xr = indices(prhs[10], "xr", &nr);
    zr = indices(prhs[11], "zr", &nr2);
    if (nr != nr2) mexErrMsgIdAndTxt("mnad:arg", "xr and zr differ in length");
    ss = as_int(scalar(prhs[12], "ss"), "ss");
    indx = as_int(scalar(prhs[13], "indx"), "indx");
    plhs[0] = mxCreateDoubleMatrix((size_t)nt, nr, mxREAL);
    if (indx == 1 && nlhs > 1) {
        plhs[1] = mxCreateDoubleMatrix((size_t)(g.nz - 2 * g.npml), (size_t)(g.nx - 2 * g.npml), mxREAL);
        illum = mxGetPr(plhs[1]);
    } else if (nlhs > 1) {
        plhs[1] = mxCreateDoubleMatrix(0, 0, mxREAL);
    }
    check(mnad_fwd(context(), &g, v, nt, w, zs, xs, (int)nr, zr, xr, ss, indx, snap_format(), 0,
                   mxGetPr(plhs[0]), NULL, NULL, NULL, illum));
    mxFree(xr);
    mxFree(zr);
}

static void cmd_adj(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    mnad_grid g;
    const double *v, *data;
    int nt, ss;
    size_t nr, nr2;
    int *xr, *zr;
    (void)nlhs;
    if (nrhs != 11)
        mexErrMsgIdAndTxt("mnad:arg", "usage: mnad_mex('adj', dx,dz,nt,dt, v, xr,zr, data, d_pml, ss)");
    g.dx = scalar(prhs[1], "dx");
    g.dz = scalar(prhs[2], "dz");
    nt = as_int(scalar(prhs[3], "nt"), "nt");
    g.dt = scalar(prhs[4], "dt");
    v = real_double(prhs[5], "v");
    g.nz = (int)mxGetM(prhs[5]);
    g.nx = (int)mxGetN(prhs[5]);
    xr = indices(prhs[6], "xr", &nr);
    zr = indices(prhs[7], "zr", &nr2);
    if (nr != nr2) mexErrMsgIdAndTxt("mnad:arg", "xr and zr differ in length");
    data = real_double(prhs[8], "data");
    if (mxGetM(prhs[8]) != (size_t)nt || mxGetN(prhs[8]) != nr)
        mexErrMsgIdAndTxt("mnad:arg", "data must be [nt x nr] = [%d x %d]", nt, (int)nr);
    g.npml = as_int(scalar(prhs[9], "d_pml"), "d_pml");
    ss = as_int(scalar(prhs[10], "ss"), "ss");
    plhs[0] = mxCreateDoubleMatrix((size_t)(g.nz - 2 * g.npml) * (size_t)(g.nx - 2 * g.npml), 1, mxREAL);
    check(mnad_born_adj(context(), &g, v, nt, (int)nr, zr, xr, data, ss, 0, mxGetPr(plhs[0])));
    mxFree(xr);
    mxFree(zr);
}

static void cmd_file(const char *cmd, int nrhs, const mxArray *prhs[])
{
    char path[4096];
    int ss;
    if (nrhs != 3 || !mxIsChar(prhs[2]))
        mexErrMsgIdAndTxt("mnad:arg", "usage: mnad_mex('%s', ss, path)", cmd);
    ss = as_int(scalar(prhs[1], "ss"), "ss");
    if (mxGetString(prhs[2], path, sizeof(path)) != 0)
        mexErrMsgIdAndTxt("mnad:arg", "path too long");
    if (strcmp(cmd, "save") == 0) check(mnad_snap_save(context(), ss, path));
    else check(mnad_snap_load(context(), ss, path));
}

/* xs, zs [ns], xr, zr [sum nrec], nrec [ns] -> a 0-based mnad_shots view
 * (arrays mxMalloc'd into *buf, freed by the caller). */
typedef struct {
    mnad_shots sh;
    int *xs, *zs, *xr, *zr, *off;
    size_t ns, ntot;
} shot_list;

static void shots_arg(const mxArray *axs, const mxArray *azs, const mxArray *axr,
                      const mxArray *azr, const mxArray *anr, shot_list *L)
{
    size_t n2, m, m2, k;
    const double *nr = real_double(anr, "nrec");
    L->xs = indices(axs, "xs", &L->ns);
    L->zs = indices(azs, "zs", &n2);
    if (L->ns != n2 || mxGetNumberOfElements(anr) != L->ns || L->ns == 0)
        mexErrMsgIdAndTxt("mnad:arg", "xs, zs and nrec must have one entry per shot");
    L->xr = indices(axr, "xr", &m);
    L->zr = indices(azr, "zr", &m2);
    if (m != m2) mexErrMsgIdAndTxt("mnad:arg", "xr and zr differ in length");
    L->off = (int *)mxMalloc((L->ns + 1) * sizeof(int));
    L->off[0] = 0;
    for (k = 0; k < L->ns; ++k) {
        const int c = as_int(nr[k], "nrec");
        if (c < 0) mexErrMsgIdAndTxt("mnad:arg", "nrec must be >= 0");
        L->off[k + 1] = L->off[k] + c;
    }
    if ((size_t)L->off[L->ns] != m)
        mexErrMsgIdAndTxt("mnad:arg", "sum(nrec) = %d differs from length(xr) = %d",
                          L->off[L->ns], (int)m);
    L->ntot = m;
    L->sh.nshots = (int)L->ns;
    L->sh.zs = L->zs;
    L->sh.xs = L->xs;
    L->sh.rec_off = L->off;
    L->sh.zr = L->zr;
    L->sh.xr = L->xr;
}

static void shots_free(shot_list *L)
{
    mxFree(L->xs);
    mxFree(L->zs);
    mxFree(L->xr);
    mxFree(L->zr);
    mxFree(L->off);
}

static const double *model_arg(const mxArray *a, const char *name, mnad_grid *g)
{
    const double *v = real_double(a, name);
    g->nz = (int)mxGetM(a);
    g->nx = (int)mxGetN(a);
    return v;
}

static void grid_args(const mxArray *const *p, mnad_grid *g)
{
    g->dx = scalar(p[0], "dx");
    g->dz = scalar(p[1], "dz");
    g->dt = scalar(p[2], "dt");
    g->npml = as_int(scalar(p[3], "d_pml"), "d_pml");
}

static const double *wavelet_arg(const mxArray *aw, const mxArray *ant, int *nt)
{
    *nt = as_int(scalar(ant, "nt"), "nt");
    if ((size_t)*nt != mxGetNumberOfElements(aw))
        mexErrMsgIdAndTxt("mnad:arg", "length(w) = %d differs from nt = %d",
                          (int)mxGetNumberOfElements(aw), *nt);
    return real_double(aw, "w");
}

static void cmd_fwd_shots(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    mnad_grid g;
    shot_list L;
    const double *v, *w;
    int nt, batch;
    (void)nlhs;
    if (nrhs != 14)
        mexErrMsgIdAndTxt("mnad:arg", "usage: mnad_mex('fwd_shots', dx,dz,dt,d_pml, v, w, nt, "
                                      "xs,zs, xr,zr, nrec, batch)");
    grid_args(prhs + 1, &g);
    v = model_arg(prhs[5], "v", &g);
    w = wavelet_arg(prhs[6], prhs[7], &nt);
    shots_arg(prhs[8], prhs[9], prhs[10], prhs[11], prhs[12], &L);
    batch = as_int(scalar(prhs[13], "batch"), "batch");
    plhs[0] = mxCreateDoubleMatrix((size_t)nt, L.ntot, mxREAL);
    check(mnad_fwd_shots(context(), &g, v, nt, w, &L.sh, batch, 0, mxGetPr(plhs[0])));
    shots_free(&L);
}

static void cmd_rtm_shots(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    mnad_grid g, gs;
    shot_list L;
    const double *v, *vs, *w;
    int nt, batch, sid = 0, ns = 0, rc;
    size_t nzi, nxi;
    double *dobs = NULL, *dsyn = NULL;
    if (nrhs != 15)
        mexErrMsgIdAndTxt("mnad:arg", "usage: mnad_mex('rtm_shots', dx,dz,dt,d_pml, v, vs, w, nt, "
                                      "xs,zs, xr,zr, nrec, batch)");
    grid_args(prhs + 1, &g);
    v = model_arg(prhs[5], "v", &g);
    gs = g;
    vs = model_arg(prhs[6], "vs", &gs);
    if (gs.nz != g.nz || gs.nx != g.nx) mexErrMsgIdAndTxt("mnad:arg", "v and vs differ in size");
    w = wavelet_arg(prhs[7], prhs[8], &nt);
    shots_arg(prhs[9], prhs[10], prhs[11], prhs[12], prhs[13], &L);
    batch = as_int(scalar(prhs[14], "batch"), "batch");
    if (g.nz <= 2 * g.npml || g.nx <= 2 * g.npml) mexErrMsgIdAndTxt("mnad:arg", "empty interior");
    nzi = (size_t)(g.nz - 2 * g.npml);
    nxi = (size_t)(g.nx - 2 * g.npml);
    plhs[0] = mxCreateDoubleMatrix(nzi, nxi, mxREAL);
    if (nlhs > 1) plhs[1] = mxCreateDoubleMatrix(nzi, nxi, mxREAL);
    if (nlhs > 2) {
        plhs[2] = mxCreateDoubleMatrix((size_t)nt, L.ntot, mxREAL);
        dobs = mxGetPr(plhs[2]);
    }
    if (nlhs > 3) {
        plhs[3] = mxCreateDoubleMatrix((size_t)nt, L.ntot, mxREAL);
        dsyn = mxGetPr(plhs[3]);
    }
    check(mnad_stack_create(context(), &g, &sid));
    rc = mnad_rtm_shots(context(), sid, &g, v, vs, nt, w, &L.sh, snap_format(), batch, 0, dobs,
                        dsyn);
    if (rc == MNAD_OK)
        rc = mnad_stack_image(context(), sid, mxGetPr(plhs[0]), nlhs > 1 ? mxGetPr(plhs[1]) : NULL,
                              &ns);
    mnad_stack_destroy(context(), sid);
    shots_free(&L);
    check(rc);
}

/* Requests in flight through the server: the output buffers a ticket writes
 * until mnad_wait returns for it (malloc'd, handed to MATLAB at 'wait'). */
typedef struct pending {
    int64_t ticket;
    double *out;
    size_t m, n;
    struct pending *next;
} pending;
static pending *g_pending = NULL;

static pending *take_pending(int64_t t)
{
    pending **pp = &g_pending;
    while (*pp && (*pp)->ticket != t) pp = &(*pp)->next;
    if (!*pp) return NULL;
    pending *p = *pp;
    *pp = p->next;
    return p;
}

static void cmd_submit(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    mnad_request q;
    mnad_grid gs;
    size_t nr, nr2;
    int *xr, *zr, nt, rc;
    int64_t t = 0;
    pending *p;
    (void)nlhs;
    if (nrhs != 15)
        mexErrMsgIdAndTxt("mnad:arg", "usage: mnad_mex('submit', kind, dx,dz,dt,d_pml, v, vs, w, nt, "
                                      "xs,zs, xr,zr, snap)");
    memset(&q, 0, sizeof(q));
    q.kind = as_int(scalar(prhs[1], "kind"), "kind");
    grid_args(prhs + 2, &q.grid);
    q.v = model_arg(prhs[6], "v", &q.grid);
    if (q.kind == MNAD_REQ_RTM) {
        gs = q.grid;
        q.v_smooth = model_arg(prhs[7], "vs", &gs);
        if (gs.nz != q.grid.nz || gs.nx != q.grid.nx)
            mexErrMsgIdAndTxt("mnad:arg", "v and vs differ in size");
    }
    q.w = wavelet_arg(prhs[8], prhs[9], &nt);
    q.nt = nt;
    q.xs = index0(scalar(prhs[10], "xs"), "xs");
    q.zs = index0(scalar(prhs[11], "zs"), "zs");
    xr = indices(prhs[12], "xr", &nr);
    zr = indices(prhs[13], "zr", &nr2);
    if (nr != nr2) mexErrMsgIdAndTxt("mnad:arg", "xr and zr differ in length");
    q.nr = (int)nr;
    q.xr = xr;
    q.zr = zr;
    q.snap_fmt = as_int(scalar(prhs[14], "snap"), "snap");
    p = (pending *)calloc(1, sizeof(*p));
    if (!p) mexErrMsgIdAndTxt("mnad:nomem", "out of host memory");
    if (q.kind == MNAD_REQ_RTM) {
        p->m = (size_t)(q.grid.nz - 2 * q.grid.npml);
        p->n = (size_t)(q.grid.nx - 2 * q.grid.npml);
    } else {
        p->m = (size_t)nt;
        p->n = nr;
    }
    p->out = (double *)calloc((p->m * p->n) > 0 ? p->m * p->n : 1, sizeof(double));
    if (!p->out) {
        free(p);
        mexErrMsgIdAndTxt("mnad:nomem", "out of host memory");
    }
    if (q.kind == MNAD_REQ_RTM) q.out_image = p->out;
    else q.out_gather = p->out;
    rc = mnad_submit(context(), &q, &t);   /* the request's inputs are copied */
    mxFree(xr);
    mxFree(zr);
    if (rc != MNAD_OK) {
        free(p->out);
        free(p);
        check(rc);
    }
    p->ticket = t;
    p->next = g_pending;
    g_pending = p;
    plhs[0] = mxCreateDoubleMatrix(1, 1, mxREAL);
    mxGetPr(plhs[0])[0] = (double)t;
}

static void cmd_wait(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    mnad_req_info info;
    int64_t t;
    int rc;
    pending *p;
    if (nrhs != 3) mexErrMsgIdAndTxt("mnad:arg", "usage: mnad_mex('wait', ticket, timeout_ms)");
    t = (int64_t)as_int(scalar(prhs[1], "ticket"), "ticket");
    memset(&info, 0, sizeof(info));
    info.status = 1;   /* filled by the call that takes the result */
    rc = mnad_wait(context(), t, as_int(scalar(prhs[2], "timeout_ms"), "timeout_ms"), &info);
    if (rc == MNAD_ERR_TIMEOUT || info.status == 1) check(rc);   /* still running / unknown */
    p = take_pending(t);
    if (rc != MNAD_OK) {
        if (p) {
            free(p->out);
            free(p);
        }
        check(rc);
    }
    if (!p) mexErrMsgIdAndTxt("mnad:state", "ticket %d was not submitted here", (int)t);
    plhs[0] = mxCreateDoubleMatrix(p->m, p->n, mxREAL);
    memcpy(mxGetPr(plhs[0]), p->out, p->m * p->n * sizeof(double));
    free(p->out);
    free(p);
    if (nlhs > 1) {
        plhs[1] = mxCreateDoubleMatrix(1, 1, mxREAL);
        mxGetPr(plhs[1])[0] = info.status;
    }
    if (nlhs > 2) {
        plhs[2] = mxCreateDoubleMatrix(1, 3, mxREAL);
        mxGetPr(plhs[2])[0] = info.submit_s;
        mxGetPr(plhs[2])[1] = info.admit_s;
        mxGetPr(plhs[2])[2] = info.done_s;
    }
}

static void cmd_serve(const char *cmd, int nrhs, const mxArray *prhs[])
{
    if (strcmp(cmd, "serve_start") == 0) {
        if (nrhs != 5)
            mexErrMsgIdAndTxt("mnad:arg", "usage: mnad_mex('serve_start', budget_bytes, "
                                          "max_active, max_points, chunk_steps)");
        check(mnad_serve_start(context(), (uint64_t)scalar(prhs[1], "budget_bytes"),
                               as_int(scalar(prhs[2], "max_active"), "max_active"),
                               (uint64_t)scalar(prhs[3], "max_points"),
                               as_int(scalar(prhs[4], "chunk_steps"), "chunk_steps")));
    } else if (strcmp(cmd, "serve_stop") == 0) {
        check(mnad_serve_stop(context()));
        while (g_pending) {   /* results of tickets nobody waited for are dropped */
            pending *p = g_pending;
            g_pending = p->next;
            free(p->out);
            free(p);
        }
    } else {   /* cancel */
        if (nrhs != 2) mexErrMsgIdAndTxt("mnad:arg", "usage: mnad_mex('cancel', ticket)");
        check(mnad_cancel(context(), (int64_t)as_int(scalar(prhs[1], "ticket"), "ticket")));
    }
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    char cmd[32];
    if (nrhs < 1 || !mxIsChar(prhs[0]) || mxGetString(prhs[0], cmd, sizeof(cmd)) != 0)
        mexErrMsgIdAndTxt("mnad:arg", "first argument must be a command string");
    if (strcmp(cmd, "version") == 0) {
        plhs[0] = mxCreateString(mnad_version());
    } else if (strcmp(cmd, "fwd") == 0) {
        cmd_fwd(nlhs, plhs, nrhs, prhs);
    } else if (strcmp(cmd, "adj") == 0) {
        cmd_adj(nlhs, plhs, nrhs, prhs);
    } else if (strcmp(cmd, "save") == 0 || strcmp(cmd, "load") == 0) {
        cmd_file(cmd, nrhs, prhs);
    } else if (strcmp(cmd, "fwd_shots") == 0) {
        cmd_fwd_shots(nlhs, plhs, nrhs, prhs);
    } else if (strcmp(cmd, "rtm_shots") == 0) {
        cmd_rtm_shots(nlhs, plhs, nrhs, prhs);
    } else if (strcmp(cmd, "submit") == 0) {
        cmd_submit(nlhs, plhs, nrhs, prhs);
    } else if (strcmp(cmd, "wait") == 0) {
        cmd_wait(nlhs, plhs, nrhs, prhs);
    } else if (strcmp(cmd, "serve_start") == 0 || strcmp(cmd, "serve_stop") == 0 ||
               strcmp(cmd, "cancel") == 0) {
        cmd_serve(cmd, nrhs, prhs);
    } else if (strcmp(cmd, "release") == 0) {
        if (nrhs != 2) mexErrMsgIdAndTxt("mnad:arg", "usage: mnad_mex('release', ss)");
        check(mnad_release(context(), as_int(scalar(prhs[1], "ss"), "ss")));
    } else {
        mexErrMsgIdAndTxt("mnad:arg", "unknown command '%s'", cmd);
    }
}
