/*
 * Helm3d_kacz_sweep_gpu_mex.c -- MEX gateway of the MI355X Kaczmarz double
 * sweep, the drop-in for Helm3d_kacz_sweep_mex, Helm3d_kacz_sweep_wnc_mex and
 * Helm3d_kacz_sweep_adj_mex (ibin/install_MEX:106-114, built from
 * tools/algorithms/3DFreqModeling/Helm3dkaczswp_mex.c), called by
 * helm3d_operto27_kaczswp.m:2-12.
 *
 *   y = Helm3d_kacz_sweep<_variant>_mex(wn, h, n, npml, omega, x, b)
 *
 * Variant flags: -DADJ (sweep on H^H), -DWNCMPLX (complex wn honoured; the
 * reference builds its _wnc gateway with a misspelt flag and ignores the
 * imaginary part).  Inputs, checks and error identifiers follow
 * Helm3dkaczswp_mex.c:123-197:
 *   n       Nsize unless one dimension of n is 3 (:148-152)
 *   x       prod(n) x P (Xsize, :157-161); real x has zero imaginary part
 *   b       prod(n) rows (Bsize, :162-165); real b has zero imaginary part
 *   y       prod(n) x P complex: column s is one forward + backward sweep of
 *           x(:,s) against b(:,s) with relaxation omega (:182-186).
 * The rows are visited in 27 colours (parallel on the GPU) instead of
 * lexicographically: the fixed point is the same, the iterates differ.
 * Checks the reference leaves to an out-of-bounds read raise errors here:
 * argument count (:nargin), n holding positive integers (:Nsize), b with a
 * column count other than P (:Bsize), numel(wn), numel(h), numel(npml)
 * (:WNsize, :Hsize, :NPMLsize).
 *
 * MATLAB build:
 *   mex -R2017b -O Helm3d_kacz_sweep_gpu_mex.c -I<repo>/include -L<lib dir> -lhelm3d \
 *       -output Helm3d_kacz_sweep_mex         (-DWNCMPLX / -DADJ for the others)
 */
#include <math.h>
#include <stdio.h>

#include "helm3d.h"
#include "mex.h"

#define ID "SLIM_release_apps:tools:algorithms:ThreeDFreqModeling:Helm3dkaczswp_mex:"

#ifdef ADJ
#define VAR_MODE HELM3D_ADJ
#else
#define VAR_MODE HELM3D_FORW
#endif
#ifdef WNCMPLX
#define VAR_WNC HELM3D_WNC
#else
#define VAR_WNC 0
#endif

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const mxArray *WN, *H, *N, *NPML, *OMEGA, *X, *B;
    const double *n;
    double numel, omega;
    double *wni = NULL;
    size_t P;
    int code, wni_alloc = 0;

    if (nrhs != 7) mexErrMsgIdAndTxt(ID "nargin", "usage: y = Helm3d_kacz_sweep_mex(wn,h,n,npml,omega,x,b)");
    if (nlhs > 1) mexErrMsgIdAndTxt(ID "nargout", "one output");
    WN = prhs[0];
    H = prhs[1];
    N = prhs[2];
    NPML = prhs[3];
    OMEGA = prhs[4];
    X = prhs[5];
    B = prhs[6];

    if (mxGetM(N) != 3 && mxGetN(N) != 3) mexErrMsgIdAndTxt(ID "Nsize", "n must be a 3-length vector");
    n = mxGetPr(N);
    for (int a = 0; a < 3; ++a)
        if (!(n[a] >= 1.0) || n[a] != floor(n[a])) mexErrMsgIdAndTxt(ID "Nsize", "n must be a 3-length vector");
    omega = mxGetScalar(OMEGA);
    numel = n[0] * n[1] * n[2];
    if ((double)mxGetM(X) != numel) mexErrMsgIdAndTxt(ID "Xsize", "x must have nx*ny*nz elements");
    P = mxGetN(X);
    if ((double)mxGetM(B) != numel) mexErrMsgIdAndTxt(ID "Bsize", "b must have nx*ny*nz elements");
    if (mxGetN(B) != P) mexErrMsgIdAndTxt(ID "Bsize", "b must have as many columns as x");
    if ((double)mxGetNumberOfElements(WN) < numel) mexErrMsgIdAndTxt(ID "WNsize", "wn must have nx*ny*nz elements");
    if (mxGetNumberOfElements(H) < 3) mexErrMsgIdAndTxt(ID "Hsize", "h must hold 3 spacings");
    if (mxGetNumberOfElements(NPML) < 6) mexErrMsgIdAndTxt(ID "NPMLsize", "npml must be a 2 x 3 matrix");

    if (VAR_WNC) {
        wni = mxGetPi(WN);
        if (!wni) {
            wni = (double *)mxCalloc((size_t)numel, sizeof(double));
            wni_alloc = 1;
        }
    }
    plhs[0] = mxCreateDoubleMatrix((mwSize)numel, (mwSize)P, mxCOMPLEX);
    code = helm3d_kaczswp(VAR_MODE | VAR_WNC, mxGetPr(WN), wni, mxGetPr(H), n, 3, mxGetPr(NPML), omega, mxGetPr(X),
                          mxGetPi(X), (long long)mxGetM(X), (long long)P, mxGetPr(B), mxGetPi(B),
                          (long long)mxGetM(B), (long long)mxGetN(B), mxGetPr(plhs[0]), mxGetPi(plhs[0]));
    if (wni_alloc) mxFree(wni);
    if (code == HELM3D_ERR_NSIZE) mexErrMsgIdAndTxt(ID "Nsize", "n must be a 3-length vector");
    if (code == HELM3D_ERR_XSIZE) mexErrMsgIdAndTxt(ID "Xsize", "x must have nx*ny*nz elements");
    if (code == HELM3D_ERR_BSIZE) mexErrMsgIdAndTxt(ID "Bsize", "b must have nx*ny*nz elements");
    if (code != HELM3D_OK) {
        char id[160];
        snprintf(id, sizeof(id), "SLIM_release_apps:tools:algorithms:ThreeDFreqModeling:Helm3d_kacz_sweep_gpu:%s",
                 code == HELM3D_ERR_HIP ? "hip" : code == HELM3D_ERR_NOMEM ? "nomem" : "failure");
        mexErrMsgIdAndTxt(id, "libhelm3d error %d: %s", code, helm3d_last_error());
    }
}
