/*
 * mgcg1_gputo.c -- the mgcg1 MEX function on libgputo.
 *   mex -R2018a mgcg1_gputo.c -I../../include -L.. -lgputo -output mgcg1
 *
 * Accepts the argument lists of both reference MEX files:
 *   CUDA   (Cantilever/TOP3D125MGCGCUDA/mgcg1.cu:120-182)
 *     [cgres,cgiters] = mgcg1(Ke0,F,U,CX,fixed',nfd,nelx,nely,nelz,nl,cgtol,cgmax[,printLev])
 *   OpenMP (Cantilever/omp/mgcg1.c:51-105)
 *     [cgres,cgiters] = mgcg1(Ke0,F,U,CX,fixed',WRK,nfd,nelx,nely,nelz,nl,cgtol,cgmax)
 * The two 13-argument forms differ in argument 6: WRK is a vector
 * (100*nDof doubles in top3D125mod.m), printLev a scalar.  Like the
 * reference, U (argument 3) is overwritten with the solution in the MATLAB
 * frame, and cgiters is the 0-based exit index.  A WRK shorter than the
 * reference's estimate gives cgiters = -1 (mgcg1.c:89-97); library failures
 * give the library's negative status.
 */
#include "gputo_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    if (nrhs != 12 && nrhs != 13) mexErrMsgTxt("mgcg1: 12 or 13 input arguments");
    const int omp = nrhs == 13 && mxGetNumberOfElements(prhs[5]) > 1;
    const int o = omp ? 1 : 0;
    const int nfd = (int)mxGetScalar(prhs[5 + o]);
    const int nelx = (int)mxGetScalar(prhs[6 + o]), nely = (int)mxGetScalar(prhs[7 + o]);
    const int nelz = (int)mxGetScalar(prhs[8 + o]), nl = (int)mxGetScalar(prhs[9 + o]);
    const double tol = mxGetScalar(prhs[10 + o]);
    const int maxit = (int)mxGetScalar(prhs[11 + o]);
    const int print_level = (!omp && nrhs == 13) ? (int)mxGetScalar(prhs[12]) : 0;
    const long long ndof = 3LL * (nelx + 1) * (nely + 1) * (nelz + 1);
    if (mxGetNumberOfElements(prhs[0]) != 576 || (long long)mxGetNumberOfElements(prhs[1]) != ndof ||
        (long long)mxGetNumberOfElements(prhs[2]) != ndof ||
        (long long)mxGetNumberOfElements(prhs[3]) != (long long)nelx * nely * nelz ||
        (long long)mxGetNumberOfElements(prhs[4]) < nfd)
        mexErrMsgTxt("mgcg1: array sizes do not match nelx, nely, nelz and nfd");
    double relres = 0.0;
    int nit = -1;
    if (omp && (long long)mxGetNumberOfElements(prhs[5]) < gputo_est_workspace(nfd, nl, nelx, nely, nelz)) {
        nit = -1;
    } else {
        gputo_handle *h = gputo_mex_handle(nelx, nely, nelz, nl, GPUTO_PRESET_REF_C, 0, print_level);
        if (!h)
            nit = GPUTO_ERR_DEVICE;
        else
            gputo_mgcg1(h, mxGetDoubles(prhs[0]), mxGetDoubles(prhs[1]), mxGetDoubles(prhs[2]), mxGetDoubles(prhs[3]),
                        mxGetDoubles(prhs[4]), nfd, tol, maxit, &relres, &nit);
    }
    plhs[0] = mxCreateDoubleScalar(relres);
    plhs[1] = mxCreateDoubleScalar((double)nit);
}
