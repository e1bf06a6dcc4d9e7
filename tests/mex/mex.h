/*
 * mex.h -- the part of MATLAB's MEX API that the libgputo shims
 * (gpu-to_amd/mex/mgcg1_gputo.c, mgcg9_gputo.c) use, implemented in
 * mex_emul.c so that the shims can be compiled and called from the tests
 * without MATLAB.  Test infrastructure only; a MATLAB build of the shims
 * uses MATLAB's own mex.h.
 */
#ifndef GPUTO_TEST_MEX_H
#define GPUTO_TEST_MEX_H

#include <stddef.h>

typedef struct mxArray_tag mxArray;

size_t mxGetNumberOfElements(const mxArray *a);
double mxGetScalar(const mxArray *a);
double *mxGetDoubles(const mxArray *a);
mxArray *mxCreateDoubleScalar(double v);
void mexErrMsgTxt(const char *msg);
int mexAtExit(void (*fn)(void));

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);

/* test harness */
mxArray *mx_wrap(double *data, size_t n); /* a double column vector over caller memory */
void mx_free(mxArray *a);
int mex_call(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]); /* 0, or -1 after mexErrMsgTxt */
const char *mex_last_error(void);
int mex_unload(void); /* runs the mexAtExit functions, as clear mex would; returns how many ran */

#endif
