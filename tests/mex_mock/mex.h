/* mex.h -- a minimal stand-in for the MATLAB MEX/mx C API, written for this
 * project's tests: declarations of the public names mnad_mex.c uses, with the
 * calling conventions MATLAB documents.  mock_mx.c implements them so the
 * gateway's argument handling and error mapping run without MATLAB. */
#ifndef MNAD_MOCK_MEX_H
#define MNAD_MOCK_MEX_H
#include <stdbool.h>
#include <stddef.h>

typedef size_t mwSize;
typedef struct mxArray_tag mxArray;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity flag);
mxArray *mxCreateString(const char *str);
double *mxGetPr(const mxArray *pa);
size_t mxGetM(const mxArray *pa);
size_t mxGetN(const mxArray *pa);
size_t mxGetNumberOfElements(const mxArray *pa);
bool mxIsDouble(const mxArray *pa);
bool mxIsComplex(const mxArray *pa);
bool mxIsSparse(const mxArray *pa);
bool mxIsChar(const mxArray *pa);
int mxGetString(const mxArray *pa, char *buf, mwSize buflen);
void *mxMalloc(size_t n);
void mxFree(void *ptr);
void mexErrMsgIdAndTxt(const char *identifier, const char *fmt, ...);
int mexAtExit(void (*exit_fcn)(void));

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);
#endif
