/* mock_mx.c -- implementation of tests/mex_mock/mex.h for the gateway tests,
 * plus mock_* helpers that build arrays and call mexFunction, catching
 * mexErrMsgIdAndTxt with setjmp/longjmp as MATLAB would unwind. */
#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

struct mxArray_tag {
    size_t m, n;
    int is_char;
    double *pr;
    char *str;
};

static jmp_buf g_jmp;
static char g_id[128], g_msg[1024];
static void (*g_exit)(void) = NULL;

mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity flag)
{
    (void)flag;
    mxArray *a = calloc(1, sizeof(*a));
    a->m = m;
    a->n = n;
    a->pr = calloc((m * n > 0) ? m * n : 1, sizeof(double));
    return a;
}

mxArray *mxCreateString(const char *s)
{
    mxArray *a = calloc(1, sizeof(*a));
    a->m = 1;
    a->n = strlen(s);
    a->is_char = 1;
    a->str = malloc(a->n + 1);
    memcpy(a->str, s, a->n + 1);
    return a;
}

double *mxGetPr(const mxArray *a) { return a->is_char ? NULL : a->pr; }
size_t mxGetM(const mxArray *a) { return a->m; }
size_t mxGetN(const mxArray *a) { return a->n; }
size_t mxGetNumberOfElements(const mxArray *a) { return a->m * a->n; }
bool mxIsDouble(const mxArray *a) { return !a->is_char; }
bool mxIsComplex(const mxArray *a) { (void)a; return false; }
bool mxIsSparse(const mxArray *a) { (void)a; return false; }
bool mxIsChar(const mxArray *a) { return a->is_char != 0; }

int mxGetString(const mxArray *a, char *buf, mwSize len)
{
    if (!a->is_char || strlen(a->str) + 1 > len) return 1;
    strcpy(buf, a->str);
    return 0;
}

void *mxMalloc(size_t n) { return malloc(n); }
void mxFree(void *p) { free(p); }

void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...)
{
    va_list ap;
    snprintf(g_id, sizeof(g_id), "%s", id);
    va_start(ap, fmt);
    vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
    va_end(ap);
    longjmp(g_jmp, 1);
}

int mexAtExit(void (*fn)(void))
{
    g_exit = fn;
    return 0;
}

/* ---- helpers for ctypes ---- */
mxArray *mock_double(size_t m, size_t n, const double *data)
{
    mxArray *a = mxCreateDoubleMatrix(m, n, mxREAL);
    if (data) memcpy(a->pr, data, m * n * sizeof(double));
    return a;
}

mxArray *mock_string(const char *s) { return mxCreateString(s); }

const double *mock_data(const mxArray *a) { return a->pr; }
const char *mock_str(const mxArray *a) { return a->str; }

void mock_free(mxArray *a)
{
    if (!a) return;
    free(a->pr);
    free(a->str);
    free(a);
}

/* Returns 0 on success, 1 if the gateway raised (id/msg then hold the error). */
int mock_call(int nlhs, mxArray **plhs, int nrhs, mxArray **prhs, char *id, char *msg)
{
    if (setjmp(g_jmp)) {
        strcpy(id, g_id);
        strcpy(msg, g_msg);
        return 1;
    }
    mexFunction(nlhs, plhs, nrhs, (const mxArray **)prhs);
    return 0;
}

void mock_unload(void)
{
    if (g_exit) g_exit();
    g_exit = NULL;
}
