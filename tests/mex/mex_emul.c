/*
 * mex_emul.c -- MEX API emulation for the shim tests (see mex.h): double
 * column vectors over caller memory, scalars, mexErrMsgTxt as a longjmp
 * back to mex_call, mexAtExit functions run by mex_unload.
 */
#include <setjmp.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

struct mxArray_tag {
    double *pr;
    size_t n;
    int owns;
};

static jmp_buf g_jmp;
static int g_in_call = 0;
static char g_err[256];
static void (*g_exit[16])(void);
static int g_nexit = 0;

size_t mxGetNumberOfElements(const mxArray *a) { return a->n; }
double mxGetScalar(const mxArray *a) { return a->n ? a->pr[0] : 0.0; }
double *mxGetDoubles(const mxArray *a) { return a->pr; }

mxArray *mxCreateDoubleScalar(double v)
{
    mxArray *a = (mxArray *)malloc(sizeof(mxArray));
    a->pr = (double *)malloc(sizeof(double));
    a->pr[0] = v;
    a->n = 1;
    a->owns = 1;
    return a;
}

void mexErrMsgTxt(const char *msg)
{
    strncpy(g_err, msg, sizeof(g_err) - 1);
    g_err[sizeof(g_err) - 1] = 0;
    if (g_in_call) longjmp(g_jmp, 1);
    abort();
}

int mexAtExit(void (*fn)(void))
{
    for (int i = 0; i < g_nexit; ++i)
        if (g_exit[i] == fn) return 0;
    if (g_nexit < 16) g_exit[g_nexit++] = fn;
    return 0;
}

mxArray *mx_wrap(double *data, size_t n)
{
    mxArray *a = (mxArray *)malloc(sizeof(mxArray));
    a->pr = data;
    a->n = n;
    a->owns = 0;
    return a;
}

void mx_free(mxArray *a)
{
    if (!a) return;
    if (a->owns) free(a->pr);
    free(a);
}

int mex_call(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    g_err[0] = 0;
    g_in_call = 1;
    if (setjmp(g_jmp)) {
        g_in_call = 0;
        return -1;
    }
    mexFunction(nlhs, plhs, nrhs, prhs);
    g_in_call = 0;
    return 0;
}

const char *mex_last_error(void) { return g_err; }

int mex_unload(void)
{
    const int n = g_nexit;
    for (int i = 0; i < n; ++i) g_exit[i]();
    g_nexit = 0;
    return n;
}
