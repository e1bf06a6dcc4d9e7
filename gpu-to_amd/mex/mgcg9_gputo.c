/*
 * mgcg9_gputo.c -- the mgcg9 MEX function on libgputo.
 *   mex -R2018a mgcg9_gputo.c -I../../include -L.. -lgputo -output mgcg9
 *
 *   [cgres,cgiters] = mgcg9(Ke0,F,U,CX,fixed',nfd,nelx,nely,nelz,nl,cgtol,cgmax,
 *                           nswp,printLev,nu,hx,hy,hz)
 * (Homogenization/QNC/mgcg9.cu:124-182, called at crosswalk.m:183 and
 * QNC_V1_mod1m.m:199).  The mgcg9 engine is libgputo's GPUTO_PRESET_MGCG9
 * with the given sweep count; nu and hx..hz are parsed and, as in the
 * reference's kernels, unused.
 */
#include "gputo_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    if (nrhs != 18) mexErrMsgTxt("mgcg9: 18 input arguments");
    const int nfd = (int)mxGetScalar(prhs[5]);
    const int nelx = (int)mxGetScalar(prhs[6]), nely = (int)mxGetScalar(prhs[7]);
    const int nelz = (int)mxGetScalar(prhs[8]), nl = (int)mxGetScalar(prhs[9]);
    const double tol = mxGetScalar(prhs[10]);
    const int maxit = (int)mxGetScalar(prhs[11]);
    const int nswp = (int)mxGetScalar(prhs[12]);
    const int print_level = (int)mxGetScalar(prhs[13]);
    const long long ndof = 3LL * (nelx + 1) * (nely + 1) * (nelz + 1);
    if (mxGetNumberOfElements(prhs[0]) != 576 || (long long)mxGetNumberOfElements(prhs[1]) != ndof ||
        (long long)mxGetNumberOfElements(prhs[2]) != ndof ||
        (long long)mxGetNumberOfElements(prhs[3]) != (long long)nelx * nely * nelz ||
        (long long)mxGetNumberOfElements(prhs[4]) < nfd)
        mexErrMsgTxt("mgcg9: array sizes do not match nelx, nely, nelz and nfd");
    double relres = 0.0;
    int nit = -1;
    gputo_handle *h = gputo_mex_handle(nelx, nely, nelz, nl, GPUTO_PRESET_MGCG9, nswp, print_level);
    if (!h)
        nit = GPUTO_ERR_DEVICE;
    else
        gputo_mgcg1(h, mxGetDoubles(prhs[0]), mxGetDoubles(prhs[1]), mxGetDoubles(prhs[2]), mxGetDoubles(prhs[3]),
                    mxGetDoubles(prhs[4]), nfd, tol, maxit, &relres, &nit);
    plhs[0] = mxCreateDoubleScalar(relres);
    plhs[1] = mxCreateDoubleScalar((double)nit);
}
