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
 * The fwd/adj argument lists are those of fwd_PDE_exact_MNAD.m:1 and
 * Born_adj_MNAD.m:1 without the nine operator matrices; grid indices are the
 * reference's 1-based padded indices (main_Acoustic_Data.m:120-126).
 * Snapshot storage: f16, or MNAD_SNAP=f32|bf16 in the environment.
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
    if (g_ctx) mnad_ctx_destroy(g_ctx);
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

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    char cmd[16];
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
    } else if (strcmp(cmd, "release") == 0) {
        if (nrhs != 2) mexErrMsgIdAndTxt("mnad:arg", "usage: mnad_mex('release', ss)");
        check(mnad_release(context(), as_int(scalar(prhs[1], "ss"), "ss")));
    } else {
        mexErrMsgIdAndTxt("mnad:arg", "unknown command '%s'", cmd);
    }
}
