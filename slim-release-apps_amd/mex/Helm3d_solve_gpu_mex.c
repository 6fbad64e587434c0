/*
 * Helm3d_solve_gpu_mex.c -- solve-level MEX gateway: x = H \ b (or H' \ b)
 * for the 27-point Helmholtz operator on the MI355X, in complex128.
 *
 *   [x, info] = Helm3d_solve_gpu_mex(wn, h, n, npml, b [, mode [, tol
 *                                    [, maxinnerit [, maxit]]]])
 *
 * It stands where linearsolve.m (:20-48, 94-97, 139-141) runs FGMRES.m on the
 * matrix-free operator with the identity preconditioner: mode 1 solves
 * H x = b, any other mode H^H x = b (:44-48), from x = 0; tol, maxinnerit
 * (restart) and maxit (restart cycles) are the LinSolveOpts fields, with
 * LinSolveOpts.m:55-60's defaults (1e-6, 5, 10000).  One column of b runs
 * FGMRES.m; several columns run FGMRES_MT.m (FGMRES.m:35-37), the lockstep
 * multi-right-hand-side form.  wn, h, n, npml are the Helm3dmvp_mex
 * arguments (complex wn selects the complex-wavenumber operator).
 *   x     prod(n) x P complex
 *   info  4 x P: iterations, restart cycles, final relative residual
 *         ||b - H x|| / ||b|| (FP64), converged (1 / 0) per column; a column
 *         that stops at maxit above tol is returned without an error, as
 *         FGMRES.m returns.
 * Errors: ...:Helm3d_solve_gpu_mex:nargin, Nsize (n), Bsize (rows of b),
 * WNsize, Hsize, NPMLsize, and ...:hip / nomem / failure from the library.
 *
 * The gateway keeps one device context for the life of the MEX file (the
 * Krylov workspace is reused from call to call) and releases it from its
 * mexAtExit handler (clear mex, MATLAB exit).  It makes no HIP calls of its
 * own: every device allocation, copy and launch is made and checked inside
 * libhelm3d (helm3d_set_model_host, helm3d_solve_host).
 *
 * MATLAB build:
 *   mex -R2017b -O Helm3d_solve_gpu_mex.c -I<repo>/include -L<lib dir> -lhelm3d
 */
#include <math.h>
#include <stdio.h>

#include "helm3d.h"
#include "mex.h"

#define ID "SLIM_release_apps:tools:algorithms:ThreeDFreqModeling:Helm3d_solve_gpu_mex:"

static helm3d_ctx *g_ctx = NULL;

static void release_ctx(void)
{
    if (g_ctx) helm3d_destroy(g_ctx);
    g_ctx = NULL;
}

static void raise_lib(int code)
{
    if (code == HELM3D_ERR_NSIZE) mexErrMsgIdAndTxt(ID "Nsize", "n must be a 3-length vector");
    if (code == HELM3D_ERR_BSIZE) mexErrMsgIdAndTxt(ID "Bsize", "b must have nx*ny*nz rows");
    if (code == HELM3D_ERR_HIP) mexErrMsgIdAndTxt(ID "hip", "libhelm3d: %s", helm3d_last_error());
    if (code == HELM3D_ERR_NOMEM) mexErrMsgIdAndTxt(ID "nomem", "libhelm3d: %s", helm3d_last_error());
    mexErrMsgIdAndTxt(ID "failure", "libhelm3d error %d: %s", code, helm3d_last_error());
}

static double opt_arg(int nrhs, const mxArray *prhs[], int k, double dflt)
{
    return (nrhs > k && mxGetNumberOfElements(prhs[k]) > 0) ? mxGetScalar(prhs[k]) : dflt;
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const mxArray *WN, *H, *N, *NPML, *B;
    const double *n;
    double numel;
    size_t P;
    int code, mode;
    helm3d_solve_opts o;
    helm3d_solve_info *info;

    if (nrhs < 5 || nrhs > 9)
        mexErrMsgIdAndTxt(ID "nargin", "usage: [x,info] = Helm3d_solve_gpu_mex(wn,h,n,npml,b,mode,tol,maxinnerit,maxit)");
    if (nlhs > 2) mexErrMsgIdAndTxt(ID "nargout", "at most two outputs");
    WN = prhs[0];
    H = prhs[1];
    N = prhs[2];
    NPML = prhs[3];
    B = prhs[4];
    if (mxGetM(N) != 3 && mxGetN(N) != 3) mexErrMsgIdAndTxt(ID "Nsize", "n must be a 3-length vector");
    n = mxGetPr(N);
    for (int a = 0; a < 3; ++a)
        if (!(n[a] >= 1.0) || n[a] != floor(n[a])) mexErrMsgIdAndTxt(ID "Nsize", "n must be a 3-length vector");
    numel = n[0] * n[1] * n[2];
    if ((double)mxGetM(B) != numel) mexErrMsgIdAndTxt(ID "Bsize", "b must have nx*ny*nz rows");
    P = mxGetN(B);
    if ((double)mxGetNumberOfElements(WN) < numel) mexErrMsgIdAndTxt(ID "WNsize", "wn must have nx*ny*nz elements");
    if (mxGetNumberOfElements(H) < 3) mexErrMsgIdAndTxt(ID "Hsize", "h must hold 3 spacings");
    if (mxGetNumberOfElements(NPML) < 6) mexErrMsgIdAndTxt(ID "NPMLsize", "npml must be a 2 x 3 matrix");

    mode = (int)opt_arg(nrhs, prhs, 5, 1.0);
    helm3d_solve_opts_default(&o);
    o.tol = opt_arg(nrhs, prhs, 6, 1e-6);
    o.maxinnerit = (int)opt_arg(nrhs, prhs, 7, 5.0);
    o.maxit = (long long)opt_arg(nrhs, prhs, 8, 10000.0);
    o.use_x0 = 0;

    if (!g_ctx) {
        code = helm3d_create(-1, NULL, &g_ctx);
        if (code != HELM3D_OK) {
            g_ctx = NULL;
            raise_lib(code);
        }
        mexAtExit(release_ctx);
    }
    code = helm3d_set_model_host(g_ctx, n, 3, mxGetPr(H), mxGetPr(NPML), mxGetPr(WN), mxGetPi(WN));
    if (code != HELM3D_OK) raise_lib(code);

    plhs[0] = mxCreateDoubleMatrix((mwSize)numel, (mwSize)P, mxCOMPLEX);
    info = (helm3d_solve_info *)mxCalloc(P ? P : 1, sizeof(helm3d_solve_info));
    code = helm3d_solve_host(g_ctx, mode, mxGetPr(B), mxGetPi(B), (long long)mxGetM(B), (long long)P,
                             mxGetPr(plhs[0]), mxGetPi(plhs[0]), &o, info);
    if (code != HELM3D_OK && code != HELM3D_ERR_NOTCONV) {
        mxFree(info);
        raise_lib(code);
    }
    if (nlhs > 1) {
        double *d;
        plhs[1] = mxCreateDoubleMatrix(4, (mwSize)P, mxREAL);
        d = mxGetPr(plhs[1]);
        for (size_t q = 0; q < P; ++q) {
            d[4 * q + 0] = (double)info[q].iters;
            d[4 * q + 1] = (double)info[q].cycles;
            d[4 * q + 2] = info[q].relres;
            d[4 * q + 3] = info[q].converged ? 1.0 : 0.0;
        }
    }
    mxFree(info);
}
