/*
 * Helm3dmvp_gpu_mex.c -- MEX gateway of the MI355X 27-point Helmholtz
 * stencil, the drop-in for the six Helm3dmvp_*_mex binaries that
 * ibin/install_MEX:88-105 builds from
 * tools/algorithms/3DFreqModeling/Helm3dmvp_mex.c.
 *
 *   y = Helm3dmvp_<variant>_mex(wn, h, n, npml, x, nthreads)
 *
 * Built once per variant with the reference's flags (-DADJ, -DDERIV,
 * -DWNCMPLX) and output names, so helm3d_operto27_mvp.m:22-42 calls it
 * unchanged.  Inputs, checks, error identifiers and the output follow
 * Helm3dmvp_mex.c:58-107:
 *   wn       prod(n) values (real; complex read only with -DWNCMPLX)
 *   h        [hx hy hz]
 *   n        3 values (Nsize unless one dimension of n is 3, :75-79)
 *   npml     2 x 3 [lo; hi] per axis
 *   x        prod(n) rows (Xsize otherwise, :81-85), first column used; a
 *            real x has a zero imaginary part (:87-93)
 *   nthreads read and ignored (the product runs on the current HIP device)
 *   y        prod(n) x 1 complex, fully written (:96-102).
 * Checks the reference leaves to an out-of-bounds read raise errors here:
 * the argument count (:nargin), n holding positive integers (:Nsize),
 * numel(wn) < prod(n) (:WNsize), numel(h) < 3 (:Hsize), numel(npml) < 6
 * (:NPMLsize).  Library failures raise ...:Helm3dmvp_gpu:<code> with the
 * library's message.
 *
 * MATLAB build (one output per variant, against libhelm3d.so):
 *   mex -R2017b -O Helm3dmvp_gpu_mex.c -I<repo>/include -L<lib dir> -lhelm3d \
 *       -output Helm3dmvp_forw_mex            (and -DADJ ... for the others)
 */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "helm3d.h"
#include "mex.h"

#define ID "SLIM_release_apps:tools:algorithms:ThreeDFreqModeling:"

#ifdef ADJ
#define VAR_MODE HELM3D_ADJ
#else
#define VAR_MODE HELM3D_FORW
#endif
#ifdef DERIV
#define VAR_DERIV HELM3D_DERIV
#else
#define VAR_DERIV 0
#endif
#ifdef WNCMPLX
#define VAR_WNC HELM3D_WNC
#else
#define VAR_WNC 0
#endif

static void raise_lib(int code)
{
    char id[160];
    const char *what = code == HELM3D_ERR_HIP     ? "hip"
                       : code == HELM3D_ERR_NOMEM ? "nomem"
                       : code == HELM3D_ERR_ARG   ? "arg"
                                                  : "failure";
    if (code == HELM3D_ERR_NSIZE) mexErrMsgIdAndTxt(ID "Helm3dmvp_mex:Nsize", "n must be a 3-length vector");
    if (code == HELM3D_ERR_XSIZE)
        mexErrMsgIdAndTxt(ID "Helm3dkaczswp_mex:Xsize", "x must have nx*ny*nz elements");
    snprintf(id, sizeof(id), ID "Helm3dmvp_gpu:%s", what);
    mexErrMsgIdAndTxt(id, "libhelm3d error %d: %s", code, helm3d_last_error());
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const mxArray *WN, *H, *N, *NPML, *X, *NTHREADS;
    const double *n;
    double numel;
    double *wni = NULL, *yr, *yi;
    const double *xi;
    int code, nthreads, wni_alloc = 0;

    if (nrhs != 6)
        mexErrMsgIdAndTxt(ID "Helm3dmvp_mex:nargin", "usage: y = Helm3dmvp_mex(wn,h,n,npml,x,nthreads)");
    if (nlhs > 1) mexErrMsgIdAndTxt(ID "Helm3dmvp_mex:nargout", "one output");
    WN = prhs[0];
    H = prhs[1];
    N = prhs[2];
    NPML = prhs[3];
    X = prhs[4];
    NTHREADS = prhs[5];

    if (mxGetM(N) != 3 && mxGetN(N) != 3)
        mexErrMsgIdAndTxt(ID "Helm3dmvp_mex:Nsize", "n must be a 3-length vector");
    n = mxGetPr(N);
    for (int a = 0; a < 3; ++a)
        if (!(n[a] >= 1.0) || n[a] != floor(n[a]))
            mexErrMsgIdAndTxt(ID "Helm3dmvp_mex:Nsize", "n must be a 3-length vector");
    numel = n[0] * n[1] * n[2];
    if ((double)mxGetM(X) != numel)
        mexErrMsgIdAndTxt(ID "Helm3dkaczswp_mex:Xsize", "x must have nx*ny*nz elements");
    if ((double)mxGetNumberOfElements(WN) < numel)
        mexErrMsgIdAndTxt(ID "Helm3dmvp_mex:WNsize", "wn must have nx*ny*nz elements");
    if (mxGetNumberOfElements(H) < 3) mexErrMsgIdAndTxt(ID "Helm3dmvp_mex:Hsize", "h must hold 3 spacings");
    if (mxGetNumberOfElements(NPML) < 6)
        mexErrMsgIdAndTxt(ID "Helm3dmvp_mex:NPMLsize", "npml must be a 2 x 3 matrix");
    nthreads = (int)mxGetScalar(NTHREADS);

    if (VAR_WNC) {
        wni = mxGetPi(WN);
        if (!wni) {  /* real wn under the complex-wavenumber variant: zero imaginary part */
            wni = (double *)mxCalloc((size_t)numel, sizeof(double));
            wni_alloc = 1;
        }
    }
    xi = mxIsComplex(X) ? mxGetPi(X) : NULL; /* NULL: zero imaginary part */

    plhs[0] = mxCreateDoubleMatrix((mwSize)numel, 1, mxCOMPLEX);
    yr = mxGetPr(plhs[0]);
    yi = mxGetPi(plhs[0]);
    code = helm3d_mvp(VAR_MODE | VAR_DERIV | VAR_WNC, mxGetPr(WN), wni, mxGetPr(H), n, 3, mxGetPr(NPML), yr, yi,
                      mxGetPr(X), xi, (long long)mxGetM(X), nthreads);
    if (wni_alloc) mxFree(wni);
    if (code != HELM3D_OK) raise_lib(code);
}
