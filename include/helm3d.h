/*
 * helm3d.h -- C ABI of the MI355X (gfx950) 3D frequency-domain Helmholtz
 * library: the 27-point mixed-grid stencil with PML (forward, adjoint,
 * derivative, complex wavenumber), its restarted-GMRES solve, mixed-precision
 * and batched solves, and the FWI gradient accumulation.
 *
 * Plain C, plain pointers and sizes.  Every entry point returns a status
 * (HELM3D_OK = 0) and never longjmps, so a MATLAB MEX gateway can map codes to
 * its error identifiers.  The library keeps no hidden global state except the
 * HIP runtime; contexts are independent and may be used from different host
 * threads (one context per thread at a time).
 *
 * Conventions (all follow SLIM-release-apps tools/algorithms/3DFreqModeling):
 *   grid      n = (nx, ny, nz) points INCLUDING the PML; x fastest,
 *             idx = i + j*nx + k*nx*ny           (Helm3d_27pt.h:21-23)
 *   npml      6 values [xlo, xhi, ylo, yhi, zlo, zhi] = MATLAB's 2x3 [lo;hi]
 *             matrix in column-major order        (Helm3dmvp_mex.c:24, 71)
 *   h         grid spacings [hx, hy, hz] in metres
 *   wn        squared wavenumber on the PML grid, (2 pi f)^2 / v^2 for m/s
 *             (param2wavenum.m:20-29)
 *   host vectors at the drop-in boundary are split real/imaginary double
 *             arrays, as mxGetPr/mxGetPi deliver them (Helm3dmvp_mex.c:67-98);
 *   device vectors are interleaved complex (re, im) of double or float.
 */
#ifndef HELM3D_H
#define HELM3D_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define HELM3D_OK          0
#define HELM3D_ERR_NSIZE   1  /* n not a 3-vector of positive integers  (MEX id ...:Helm3dmvp_mex:Nsize)   */
#define HELM3D_ERR_XSIZE   2  /* x rows != prod(n)                        (MEX id ...:Helm3dkaczswp_mex:Xsize) */
#define HELM3D_ERR_ARG     3  /* invalid argument                                                        */
#define HELM3D_ERR_HIP     4  /* HIP runtime failure                                                     */
#define HELM3D_ERR_NOMODEL 5  /* context has no model set                                                */
#define HELM3D_ERR_NOMEM   6  /* device memory exhausted                                                 */
#define HELM3D_ERR_NOTCONV 7  /* solve stopped at maxit above tolerance; x holds the last iterate        */
#define HELM3D_ERR_BSIZE   8  /* b does not have prod(n) rows (Helm3dkaczswp_mex:Bsize)  */

/* ---- operator variants -------------------------------------------------- *
 * The six MEX binaries of ibin/install_MEX:88-105 map to
 *   Helm3dmvp_forw_mex        HELM3D_FORW
 *   Helm3dmvp_adj_mex         HELM3D_ADJ
 *   Helm3dmvp_forw_deriv_mex  HELM3D_FORW | HELM3D_DERIV
 *   Helm3dmvp_adj_deriv_mex   HELM3D_ADJ  | HELM3D_DERIV
 *   Helm3dmvp_forw_wnc_mex    HELM3D_FORW | HELM3D_WNC
 *   Helm3dmvp_adj_wnc_mex     HELM3D_ADJ  | HELM3D_WNC
 * (the dispatch of helm3d_operto27_mvp.m:22-42). */
#define HELM3D_FORW  0
#define HELM3D_ADJ   1
#define HELM3D_DERIV 2
#define HELM3D_WNC   4

/* ---- precisions --------------------------------------------------------- */
#define HELM3D_C128 0  /* complex128 vectors, FP64 wn and PML tables */
#define HELM3D_C64  1  /* complex64 vectors, FP32 wn and PML tables  */

/* =========================================================================
 * 1. Drop-in matvec at the MEX boundary (host split-complex arrays)
 * ========================================================================= */

/* Replaces do_Hmvp (Helm3d_27pt.h:827) for one compile-time variant and the
 * size checks of the gateway (Helm3dmvp_mex.c:75-85).  Synchronous: uploads
 * wn and x to the current HIP device, applies the stencil in complex128 and
 * downloads y.  y is fully overwritten (Helm3dmvp_mex.c:96-102).
 *   wni     imaginary part of wn; read only for HELM3D_WNC (may be NULL otherwise)
 *   n_len   number of entries of the MATLAB n argument (must be 3)
 *   x_rows  rows of the MATLAB x argument (must equal prod(n))
 *   xi      imaginary part of x, NULL for a real x (Helm3dmvp_mex.c:87-93)
 *   nthreads accepted for signature compatibility; ignored on the GPU. */
int helm3d_mvp(int variant, const double *wnr, const double *wni, const double *h,
               const double *n, long long n_len, const double *npml,
               double *yr, double *yi, const double *xr, const double *xi,
               long long x_rows, int nthreads);

/* do_Hmvp with its exact argument list (Helm3d_27pt.h:827), one symbol per
 * MEX variant (ibin/install_MEX:88-105); returns a status instead of void. */
int helm3d_do_Hmvp_forw(const double *wnr, const double *wni, const double *h, const double *n,
                        const double *npml, double *yr, double *yi, const double *xr,
                        const double *xi, int nthreads);
int helm3d_do_Hmvp_adj(const double *wnr, const double *wni, const double *h, const double *n,
                       const double *npml, double *yr, double *yi, const double *xr,
                       const double *xi, int nthreads);
int helm3d_do_Hmvp_forw_deriv(const double *wnr, const double *wni, const double *h,
                              const double *n, const double *npml, double *yr, double *yi,
                              const double *xr, const double *xi, int nthreads);
int helm3d_do_Hmvp_adj_deriv(const double *wnr, const double *wni, const double *h,
                             const double *n, const double *npml, double *yr, double *yi,
                             const double *xr, const double *xi, int nthreads);
int helm3d_do_Hmvp_forw_wnc(const double *wnr, const double *wni, const double *h,
                            const double *n, const double *npml, double *yr, double *yi,
                            const double *xr, const double *xi, int nthreads);
int helm3d_do_Hmvp_adj_wnc(const double *wnr, const double *wni, const double *h,
                           const double *n, const double *npml, double *yr, double *yi,
                           const double *xr, const double *xi, int nthreads);

/* =========================================================================
 * 2. Device-resident operator context
 * ========================================================================= */

typedef struct helm3d_ctx helm3d_ctx;

/* Create a context on HIP device `device` (-1: the current device).  Work is
 * issued on `stream` (a hipStream_t, NULL: the default stream). */
int helm3d_create(int device, void *stream, helm3d_ctx **out);
int helm3d_destroy(helm3d_ctx *ctx);
int helm3d_set_stream(helm3d_ctx *ctx, void *stream);

/* Set the operator H(wn) of discrete_helmholtz.m:240-251: grid, spacing, PML
 * counts and wn on the PML grid.  wn_dev points to DEVICE memory: N doubles
 * when wn_complex == 0, N interleaved complex doubles when wn_complex != 0
 * (complex frequency / WNCMPLX).  The context keeps its own FP64 and FP32
 * copies, so the caller may free wn_dev afterwards.  A model on the same grid
 * as the previous one keeps the stencil launch shape (helm3d_set_launch,
 * helm3d_autotune); a new grid starts from the default shape. */
int helm3d_set_model(helm3d_ctx *ctx, const int n[3], const double h[3], const int npml[6],
                     const void *wn_dev, int wn_complex);

/* Number of grid points of the current model (0 when none). */
long long helm3d_numel(const helm3d_ctx *ctx);

/* y = H x, H^H x (variant HELM3D_FORW / HELM3D_ADJ, optionally | HELM3D_WNC
 * ignored: the model decides) on device vectors of the given precision.
 * The matvec of helm3d_operto27_mvp.m:22-42 without host copies. */
int helm3d_apply(helm3d_ctx *ctx, int variant, int precision, const void *x_dev, void *y_dev);

/* y = (dH/dwn . dwn) x or its adjoint: the DERIV variants
 * (Helm3d_27pt.h:202-206) with a caller-supplied real dwn (device, N values
 * of the given precision) -- e.g. (2 pi f)^2 for s^2/m^2 (param2wavenum.m:31-34). */
int helm3d_apply_deriv(helm3d_ctx *ctx, int adjoint, int precision, const void *dwn_dev,
                       const void *x_dev, void *y_dev);

/* y = H x (variant HELM3D_FORW) or H^H x (HELM3D_ADJ) in complex128 for the
 * operator given by n (points per axis, PML included), h, npml[6] and wn on
 * the device (N doubles, or N interleaved complex doubles when wn_complex),
 * without any device allocation: the variant of helm3d_apply for callers that
 * cannot call hipMalloc (code under HIP graph capture, tight serving loops).
 * No counterpart in the reference.  The kernels' scratch (the PML tables and
 * the stencil's write-only sink) is a static device array of the library,
 * 1.4 MiB per GPU, held until the process ends; each axis may hold at most
 * HELM3D_NOALLOC_MAX_AXIS points (HELM3D_ERR_ARG beyond).  Work is issued on
 * `stream` (a hipStream_t, NULL: the default stream) and is asynchronous.
 * Calls that overlap on one GPU race on that array; order them (one stream,
 * or events). */
#define HELM3D_NOALLOC_MAX_AXIS 4094
int helm3d_apply_noalloc(int variant, const int n[3], const double h[3], const int npml[6],
                         const void *wn_dev, int wn_complex, const void *x_dev, void *y_dev,
                         void *stream);

/* Stencil launch shape: z planes per block (0: automatic), column-tile rows
 * (4 or 8), planes of loads in flight (1 to 4; the batched kernels use at
 * most 2) and rows per thread (1, or 2 with 8 tile rows: 256 threads per
 * 64 x 8 tile instead of 512, so two tiles share a compute unit).  The
 * result is identical for every shape; only speed changes.  Chosen per grid
 * by the autotuner (helm3d_autotune). */
int helm3d_set_launch(helm3d_ctx *ctx, int kchunk, int tile_rows, int prefetch, int rows_per_thread);
int helm3d_get_launch(helm3d_ctx *ctx, int *kchunk, int *tile_rows, int *prefetch, int *rows_per_thread,
                      int *blocks);
/* Autotuner of the stencil launch shape for this context's grid: times every
 * candidate shape (64 x 4 and 64 x 8 column tiles, the latter with one or two
 * rows per thread, prefetch {1, 2, 3, 4}, z chunks of the whole column down
 * to 8 planes, and of 4, 2 and 1 planes on grids of fewer than 128 64 x 4
 * column tiles) on the operator of `mode` (1: H, else H^H) in `precision`,
 * `reps` launches each with HIP events on the context's stream, and keeps the
 * fastest (results are identical for every shape).  No counterpart in the
 * reference; the batched solves use each item's tuned z chunk and the first
 * item's tile rows, rows per thread and prefetch. */
int helm3d_autotune(helm3d_ctx *ctx, int mode, int precision, int reps, int *kchunk, int *tile_rows,
                    int *prefetch, int *rows_per_thread, double *best_ms);

/* Context options.  HELM3D_OPT_GRAPHS (default 1): replay fixed kernel
 * sequences from captured HIP graphs -- the fixed-length inner solve of
 * helm3d_solve_mixed, and whole Arnoldi cycles of the unpreconditioned GMRES
 * on grids of at most 2M points (coarse multigrid levels); 0 launches them one
 * by one (identical results). */
#define HELM3D_OPT_GRAPHS 1
int helm3d_set_option(helm3d_ctx *ctx, int option, long long value);

/* ---- solves ------------------------------------------------------------- */

/* Options of LinSolveOpts.m:55-62 / FGMRES.m:42-45 plus the mixed-precision
 * controls.  Defaults via helm3d_solve_opts_default. */
typedef struct {
    double tol;         /* relative residual tolerance ||b - A x|| / ||b - A x0|| (1e-6) */
    int    maxinnerit;  /* restart length m (20)                                       */
    long long maxit;    /* maximum number of restart cycles (10000)                    */
    int    use_x0;      /* 0: start from x = 0 (linearsolve.m:21-23); 1: x holds x0    */
    /* mixed precision (helm3d_solve_mixed) */
    double inner_tol;   /* relative tolerance of each complex64 inner solve (1e-2);
                           0 with BiCGStab: exactly inner_maxit iterations per outer step,
                           replayed as one captured HIP graph (no host round trips)      */
    int    inner_m;     /* restart length of the inner complex64 GMRES (20)            */
    int    inner_maxit; /* max inner iterations per outer step (200)                   */
    int    inner_solver;/* HELM3D_INNER_BICGSTAB (default) or HELM3D_INNER_GMRES        */
} helm3d_solve_opts;

#define HELM3D_INNER_BICGSTAB 0  /* BICGSTAB.m as preconditioner (linearsolve.m:76-77) */
#define HELM3D_INNER_GMRES    1  /* FGMRES.m(inner_m) as preconditioner (linearsolve.m:81-82) */

typedef struct {
    long long iters;      /* Arnoldi steps (matvecs of the inner loop) of the FP64 loop  */
    long long inner_iters;/* complex64 matvecs (mixed precision only)                    */
    long long cycles;     /* restart cycles                                              */
    double relres;        /* final true relative residual ||b - A x|| / ||r0||, FP64     */
    double quasi_relres;  /* last quasi-residual (FGMRES.m:111)                          */
    int converged;        /* relres < tol                                                */
} helm3d_solve_info;

void helm3d_solve_opts_default(helm3d_solve_opts *o);

/* Restarted GMRES of FGMRES.m:33-128 with the identity preconditioner
 * (linearsolve.m:94-97, 139-141): mode 1 solves H x = b, any other mode
 * H^H x = b (linearsolve.m:44-48).  b, x: device complex128 vectors.
 * hst (optional, host) receives up to hst_cap quasi-residuals (first = 1). */
int helm3d_solve(helm3d_ctx *ctx, int mode, const void *b_dev, void *x_dev,
                 const helm3d_solve_opts *opts, helm3d_solve_info *info,
                 double *hst, long long hst_cap, long long *n_hst);

/* Mixed precision: FP64 flexible GMRES (FGMRES.m with a preconditioner,
 * :96-98, 118) whose preconditioner is a complex64 GMRES solve of the same
 * operator to inner_tol; FP64 outer residuals, FP64 Hessenberg least squares.
 * Stops on the FP64 true residual at tol. */
int helm3d_solve_mixed(helm3d_ctx *ctx, int mode, const void *b_dev, void *x_dev,
                       const helm3d_solve_opts *opts, helm3d_solve_info *info);

/* BiCGStab of BICGSTAB.m:37-74 in complex128 (LinSolveOpts.SOLVE_BICGSTAB,
 * linearsolve.m:133-136), identity preconditioner, x0 = 0.  Stops on the
 * recursive residual ||r|| / ||b|| < tol or after maxit iterations; info->relres
 * is the true FP64 residual ||b - A x|| / ||b||.  beta uses the textbook
 * rho_new / rho_old (BICGSTAB.m:59 divides a dot product by itself). */
int helm3d_solve_bicgstab(helm3d_ctx *ctx, int mode, const void *b_dev, void *x_dev,
                          const helm3d_solve_opts *opts, helm3d_solve_info *info);

/* ---- batched solves --------------------------------------------------------- */

/* Many right-hand sides, possibly on different operators and grids, solved
 * together: every kernel type runs as one launch over all items per Krylov
 * step (the stencil from a device work list of (operator, vector) items), the
 * host synchronises once per step for the whole batch, and the complex64
 * inner iterations of the mixed method are replayed from one captured HIP
 * graph covering all items.  No counterpart in the MEX layer; the plain
 * method follows FGMRES_MT.m:40-135 (the multi-RHS branch of FGMRES.m:35-37):
 * lockstep cycles, per-item Hessenberg least squares, the step loop runs
 * while the largest quasi-residual is above tol and the cycle loop while the
 * largest true residual is. */
typedef struct helm3d_batch helm3d_batch;

#define HELM3D_BATCH_PLAIN 0  /* complex128 GMRES(maxinnerit), FGMRES_MT semantics          */
#define HELM3D_BATCH_MAX_ITEMS 1024  /* items per helm3d_batch_solve call */
#define HELM3D_BATCH_MIXED 1  /* FP64 flexible GMRES + inner_maxit complex64 BiCGStab        */
                              /* iterations per outer step; items stop individually        */

int helm3d_batch_create(int device, void *stream, helm3d_batch **out);
int helm3d_batch_destroy(helm3d_batch *batch);
int helm3d_batch_set_stream(helm3d_batch *batch, void *stream);
/* Item i solves A_i x_i = b_i with A_i = H(ctx[i]) (mode 1) or H(ctx[i])^H
 * (otherwise); b[i], x[i]: device complex128 vectors of ctx[i]'s size; items
 * may share a context.  Every context must hold a model (all with real wn,
 * or all with complex wn) on the batch's device.  x starts from 0
 * (opts->use_x0 must be 0).  info: n entries (may be NULL).  Returns
 * HELM3D_ERR_NOTCONV when some item ended above tol. */
int helm3d_batch_solve(helm3d_batch *batch, int n, helm3d_ctx *const *ctx, int mode,
                       const void *const *b, void *const *x, int method,
                       const helm3d_solve_opts *opts, helm3d_solve_info *info);
/* Device workspace one item on ctx adds to a batch solve (Krylov bases,
 * inner vectors, partial sums), for callers that split a large set of
 * right-hand sides into batches that fit the device.  No counterpart in the
 * reference. */
int helm3d_batch_item_bytes(const helm3d_ctx *ctx, int method, int maxinnerit, size_t *bytes);

/* ---- Kaczmarz sweeps and CGMN ----------------------------------------------- */

/* nsweeps double Kaczmarz sweeps of x (in place) for A x = b, A = H (mode 1)
 * or H^H (otherwise), relaxation omega: the row projections of
 * Helm3dkaczswp_mex.c:60-117, forward then backward, with the rows ordered
 * by 27 colours (i mod 3, j mod 3, k mod 3) instead of lexicographically so
 * that each colour is one parallel launch (the iterates differ from the
 * lexicographic sweep; its fixed point does not).  b may be NULL (zero). */
int helm3d_kacz_sweep(helm3d_ctx *ctx, int mode, double omega, int nsweeps, void *x_dev,
                      const void *b_dev);

/* Drop-in of y = Helm3d_kacz_sweep{,_wnc,_adj}_mex(wn,h,n,npml,w,x,b)
 * (helm3d_operto27_kaczswp.m:2-12, Helm3dkaczswp_mex.c:123-197) on host
 * split-complex arrays: variant HELM3D_FORW, HELM3D_ADJ, or either with
 * HELM3D_WNC (complex wn honoured); xi, bi may be NULL (zero imaginary part).
 * x, b and y hold x_cols columns of x_rows values (column-major, as
 * mxGetPr delivers an N x P matrix); column s of y is one double sweep of
 * column s of x against column s of b (Helm3dkaczswp_mex.c:159,182-186).
 * Errors: HELM3D_ERR_NSIZE, HELM3D_ERR_XSIZE (x rows), HELM3D_ERR_BSIZE
 * (b rows, or b columns != x columns) as the MEX. */
int helm3d_kaczswp(int variant, const double *wnr, const double *wni, const double *h,
                   const double *n, long long n_len, const double *npml, double omega,
                   const double *xr, const double *xi, long long x_rows, long long x_cols,
                   const double *br, const double *bi, long long b_rows, long long b_cols,
                   double *yr, double *yi);

/* CGMN (CGMN.m) on the Kaczmarz system of linearsolve.m:52-58 with
 * omega = 1.5 and one double sweep: b' = SWP(0, b), A' y = y - SWP(y, 0), CG
 * from x = 0 until ||r'|| / ||b'|| <= tol or maxit iterations.  info->iters:
 * CG iterations; info->quasi_relres: the preconditioned residual;
 * info->relres: the true ||b - A x|| / ||b||. */
int helm3d_solve_cgmn(helm3d_ctx *ctx, int mode, const void *b_dev, void *x_dev,
                      const helm3d_solve_opts *opts, helm3d_solve_info *info);
/* CRMN (CRMN.m:46-108, conjugate residuals) on the same Kaczmarz system
 * (linearsolve.m:52-63, solver CRMN with the KACZSWP preconditioner): two
 * inner products and one double sweep per iteration; info as for CGMN. */
int helm3d_solve_crmn(helm3d_ctx *ctx, int mode, const void *b_dev, void *x_dev,
                      const helm3d_solve_opts *opts, helm3d_solve_info *info);

/* ---- host-array solves (solve-level gateways) ------------------------------ */

/* Load H(wn) into ctx from MATLAB-style double arrays: n (n_len entries, must
 * be 3 positive integers, else HELM3D_ERR_NSIZE), h[3], npml[6] as
 * helm3d_mvp takes them, and wn as split real / imaginary host arrays (wni
 * NULL: real wn).  Synchronous.  The device-memory form is helm3d_set_model;
 * this one lets a MEX gateway keep a context without HIP calls of its own. */
int helm3d_set_model_host(helm3d_ctx *ctx, const double *n, long long n_len, const double *h,
                          const double *npml, const double *wnr, const double *wni);

/* x = H \ b (mode 1) or H^H \ b (otherwise) for b_cols right-hand sides held
 * as split real / imaginary host arrays of b_rows x b_cols values
 * (column-major; bi NULL: real b), from x = 0: linearsolve.m:20-48 with the
 * FGMRES solver of LinSolveOpts (tol, maxinnerit, maxit from opts).  One
 * column runs FGMRES.m (helm3d_solve); several columns run FGMRES_MT.m
 * (FGMRES.m:35-37; helm3d_batch_solve, HELM3D_BATCH_PLAIN).  xr, xi receive
 * b_rows x b_cols values; info (may be NULL) b_cols entries.  Errors:
 * HELM3D_ERR_BSIZE when b_rows != prod(n) of the model, HELM3D_ERR_NOMODEL;
 * HELM3D_ERR_NOTCONV when a column ended above tol (x is still written). */
int helm3d_solve_host(helm3d_ctx *ctx, int mode, const double *br, const double *bi,
                      long long b_rows, long long b_cols, double *xr, double *xi,
                      const helm3d_solve_opts *opts, helm3d_solve_info *info);

/* ---- multigrid-preconditioned FGMRES ------------------------------------------ */

/* The ML_GMRES preconditioner (ML_GMRES.m:34-72, opMultigrid.m:71-110,
 * construct_multigrid.m, GMG_V.m) over device-resident levels.  ops[l] and
 * smoothers[l] are two contexts holding the same level-l operator (level 0:
 * the operator to solve; level l: the model coarsened to 2^l spacing with
 * ceil(npml / 2^l) PML points, built by the caller); the transfers are the
 * trilinear interpolation between consecutive PML-extended grids over
 * linspace(0,1,n) per axis (fine2coarse.m:40-57, LinInterp1D.m) and 1/8 of its
 * transpose.  All contexts are switched to the first one's stream. */
typedef struct helm3d_mg helm3d_mg;
int helm3d_mg_create(int nlevels, helm3d_ctx *const *ops, helm3d_ctx *const *smoothers,
                     helm3d_mg **out);
int helm3d_mg_destroy(helm3d_mg *mg);
/* Issue the hierarchy's work (transfers, solves, every level context) on
 * `stream` (a hipStream_t, NULL: the default stream) from now on; callers
 * that allocate b and x on their own stream bind it before each call, as
 * helm3d_set_stream does for one context.  No counterpart in the reference. */
int helm3d_mg_set_stream(helm3d_mg *mg, void *stream);
/* direction 0: out(level) += P in(level+1); 1: out(level+1) = P^T in(level) / 8 */
int helm3d_mg_transfer(helm3d_mg *mg, int level, int direction, const void *in_dev, void *out_dev);
/* FGMRES (FGMRES.m, flexible, opts->maxinnerit restart) on level 0
 * preconditioned by one recursive V-cycle per step: smoother FGMRES(k[1]) x
 * k[0] cycles (tol 1e-6), coarse solves FGMRES(k[3]) x k[2] cycles (tol 0.5)
 * preconditioned by the next level's V-cycle, the coarsest by another such
 * solve (k = NULL: 3, 5, 3, 5 as ML_GMRES.m:3-6). */
int helm3d_solve_mg(helm3d_mg *mg, int mode, const void *b_dev, void *x_dev,
                    const helm3d_solve_opts *opts, const int *k, helm3d_solve_info *info);
/* Mixed precision form of helm3d_solve_mg: the outer FGMRES, its basis, the
 * Hessenberg least squares and the true residuals in FP64; the V-cycle
 * preconditioner (smoothers, residuals, transfers, coarse solves) in
 * complex64 with the FP32 copy of every level's operator.  b, x complex128. */
int helm3d_solve_mg_mixed(helm3d_mg *mg, int mode, const void *b_dev, void *x_dev,
                          const helm3d_solve_opts *opts, const int *k, helm3d_solve_info *info);

/* ---- FWI gradient ------------------------------------------------------- */

#define HELM3D_GRID_PHYS 0  /* g on the physical grid: comp2phys = truncation  */
#define HELM3D_GRID_PML  1  /* g on the whole PML-extended grid               */

/* g += scale * Re( conj(U) .* (dH^H V) ): the gradient term sum_srcs(T(U)'*V)
 * of PDEfunc.m:393-394 with T_adj(z,u) = conj(u).*(dH'*z)
 * (discrete_helmholtz.m:288) and dH the DERIV stencil with real dwn (device,
 * N doubles; NULL means dwn = 1, so that a constant dwn -- s2/m2 units -- is
 * passed as scale).  U, V: device complex128 vectors on the PML grid.
 * grid HELM3D_GRID_PHYS: g holds the physical grid (n - npml_lo - npml_hi per
 * axis) and receives the truncated term, i.e. comp2phys = Ppad' of
 * discrete_helmholtz.m:231-233 (helm_cut_pml, the default).
 * grid HELM3D_GRID_PML: g holds the extended grid (N doubles); fold it with
 * helm3d_extension_adjoint for comp2phys = Pext' (:234-235).  One stencil pass:
 * reads V, U, g (and dwn), writes g. */
int helm3d_fwi_grad_accum(helm3d_ctx *ctx, const double *dwn_dev, const void *U_dev,
                          const void *V_dev, double scale, int grid, double *g_dev);

/* ---- model set-up and source / receiver operators -------------------------- */

#define HELM3D_UNIT_MS    0  /* 'm/s'    */
#define HELM3D_UNIT_S2M2  1  /* 's2/m2'  */
#define HELM3D_UNIT_S2KM2 2  /* 's2/km2' */

/* The model v (device, physical grid n_phys, x fastest) extended into the PML
 * by edge replication (Pext = phys2comp, discrete_helmholtz.m:214-217,240) and
 * mapped to the wavenumber parameter and its derivatives,
 * [wn, dwn, ddwn] = param2wavenum(Pext v, freq, unit) (param2wavenum.m:20-49).
 * Outputs: device doubles on the extended grid; dwn and ddwn may be NULL.
 * Runs on the context's stream and needs no model. */
int helm3d_param2wavenum(helm3d_ctx *ctx, const int n_phys[3], const int npml[6],
                         const double *v_dev, int unit, double freq, double *wn_dev,
                         double *dwn_dev, double *ddwn_dev);

/* g = Pext' gpml (accumulate: g += Pext' gpml): the adjoint of the edge
 * replication, opExtension.m:97-101 composed per axis.  gpml: extended grid,
 * g: physical grid (device doubles). */
int helm3d_extension_adjoint(helm3d_ctx *ctx, const int n_phys[3], const int npml[6],
                             const double *gpml_dev, double *g_dev, int accumulate);

/* Row-compressed real-weight operator on complex128 device vectors, the form
 * of the source injection Ps and receiver sampling Pr operators
 * (DataGrid.m:22-44, opInterp.m:34-57):
 *   out[oidx ? oidx[r] : r] (accumulate ? += : =) alpha * sum_e w[e] in[col[e]],
 *   e in [rowptr[r], rowptr[r+1]),  r < nrows.
 * All index and weight arrays are device memory; oidx rows must be distinct. */
int helm3d_csr_apply(helm3d_ctx *ctx, int nrows, const long long *rowptr_dev,
                     const long long *col_dev, const double *w_dev, const long long *oidx_dev,
                     double alpha, const void *in_dev, void *out_dev, int accumulate);

/* ---- in-library kernel timing ------------------------------------------ *
 * When enabled, every kernel the context launches is bracketed by HIP events
 * on its stream and the elapsed times are summed per kernel class.  Enabling
 * (or disabling) resets the totals. */
#define HELM3D_PROF_STENCIL 0  /* the 27-point stencil (matvec, residual)   */
#define HELM3D_PROF_DOTS    1  /* Gram-Schmidt dot products                 */
#define HELM3D_PROF_UPDATE  2  /* Gram-Schmidt update + norm                */
#define HELM3D_PROF_COMBINE 3  /* solution update x += V y                  */
#define HELM3D_PROF_REDUCE  4  /* second stage of the reductions            */
#define HELM3D_PROF_KINDS   5
int helm3d_profile(helm3d_ctx *ctx, int enable);
int helm3d_profile_read(helm3d_ctx *ctx, int kind, double *total_ms, long long *launches);

/* Library version string. */
const char *helm3d_version(void);

/* Text of the last error raised on the calling host thread ("" if none). */
const char *helm3d_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* HELM3D_H */
