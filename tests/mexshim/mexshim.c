/*
 * mexshim.c -- implementation of the minimal mx / mex API of matrix.h and
 * mex.h, plus a harness the tests drive through ctypes.  Test infrastructure.
 *
 * Semantics kept from MATLAB that the gateways depend on:
 *   - double matrices with separate real / imaginary storage; mxGetPi is NULL
 *     for a real array; mxCreateDoubleMatrix zero-fills;
 *   - mexErrMsgIdAndTxt does not return: it jumps back to the harness, which
 *     records the identifier and message;
 *   - memory from mxCalloc that a gateway did not mxFree is released when the
 *     call ends (MATLAB's memory manager does this); the harness counts such
 *     blocks so a test can tell a clean gateway from one that relies on it;
 *   - arrays a gateway created and did not hand out through plhs are counted
 *     the same way (MATLAB destroys them);
 *   - mexAtExit handlers run when the harness unloads the gateway.
 */
#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

struct mxArray_tag {
    size_t m, n;
    double *pr, *pi;
    struct mxArray_tag *next_live;
};

#define MAX_BLOCKS 4096
static void *g_blocks[MAX_BLOCKS];  /* live mxCalloc blocks of the current call */
static int g_nblocks = 0;
static int g_auto_freed = 0;        /* mxCalloc blocks left by the last call */
static struct mxArray_tag *g_created = NULL;  /* arrays created during the call */
static int g_arrays_dropped = 0;
static jmp_buf g_jmp;
static int g_in_call = 0;
static char g_err_id[256];
static char g_err_msg[512];
static void (*g_atexit)(void) = NULL;

double *mxGetPr(const mxArray *a) { return a ? a->pr : NULL; }
double *mxGetPi(const mxArray *a) { return a ? a->pi : NULL; }
size_t mxGetM(const mxArray *a) { return a ? a->m : 0; }
size_t mxGetN(const mxArray *a) { return a ? a->n : 0; }
size_t mxGetNumberOfElements(const mxArray *a) { return a ? a->m * a->n : 0; }
int mxIsComplex(const mxArray *a) { return a && a->pi != NULL; }
int mxIsDouble(const mxArray *a) { return a != NULL; }
double mxGetScalar(const mxArray *a) { return (a && a->m * a->n > 0) ? a->pr[0] : 0.0; }

static mxArray *new_array(size_t m, size_t n, int cplx)
{
    mxArray *a = (mxArray *)calloc(1, sizeof(mxArray));
    if (!a) return NULL;
    a->m = m;
    a->n = n;
    size_t cnt = (m * n) > 0 ? m * n : 1;
    a->pr = (double *)calloc(cnt, sizeof(double));
    if (cplx) a->pi = (double *)calloc(cnt, sizeof(double));
    return a;
}

static void free_array(mxArray *a)
{
    if (!a) return;
    free(a->pr);
    free(a->pi);
    free(a);
}

mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c)
{
    mxArray *a = new_array(m, n, c == mxCOMPLEX);
    if (a && g_in_call) {
        a->next_live = g_created;
        g_created = a;
    }
    return a;
}

void mxDestroyArray(mxArray *a)
{
    struct mxArray_tag **p = &g_created;
    while (*p) {
        if (*p == a) {
            *p = a->next_live;
            break;
        }
        p = &(*p)->next_live;
    }
    free_array(a);
}

void *mxCalloc(size_t n, size_t size)
{
    void *p = calloc(n ? n : 1, size ? size : 1);
    if (p && g_nblocks < MAX_BLOCKS) g_blocks[g_nblocks++] = p;
    return p;
}

void mxFree(void *p)
{
    for (int i = 0; i < g_nblocks; ++i)
        if (g_blocks[i] == p) {
            g_blocks[i] = g_blocks[--g_nblocks];
            break;
        }
    free(p);
}

void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...)
{
    va_list ap;
    snprintf(g_err_id, sizeof(g_err_id), "%s", id ? id : "");
    va_start(ap, fmt);
    vsnprintf(g_err_msg, sizeof(g_err_msg), fmt ? fmt : "", ap);
    va_end(ap);
    if (g_in_call) longjmp(g_jmp, 1);
    fprintf(stderr, "mexErrMsgIdAndTxt outside a call: %s %s\n", g_err_id, g_err_msg);
    abort();
}

int mexAtExit(void (*fn)(void))
{
    g_atexit = fn;
    return 0;
}

int mexPrintf(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    int r = vprintf(fmt, ap);
    va_end(ap);
    return r;
}

/* ---------------------------------------------------------------- harness */

mxArray *shim_matrix(size_t m, size_t n, int cplx) { return new_array(m, n, cplx); }
double *shim_pr(mxArray *a) { return a->pr; }
double *shim_pi(mxArray *a) { return a->pi; }
size_t shim_m(mxArray *a) { return a->m; }
size_t shim_n(mxArray *a) { return a->n; }
void shim_free(mxArray *a) { free_array(a); }
const char *shim_err_id(void) { return g_err_id; }
const char *shim_err_msg(void) { return g_err_msg; }
int shim_auto_freed(void) { return g_auto_freed; }
int shim_arrays_dropped(void) { return g_arrays_dropped; }

/* Call mexFunction; 0 on return, 1 when it raised an error (identifier and
 * message in shim_err_id / shim_err_msg).  Outputs it created are returned in
 * plhs (room for max(nlhs, 1); the caller frees them with shim_free); on an
 * error they are released here, as MATLAB does. */
int shim_call(int nlhs, mxArray **plhs, int nrhs, mxArray **prhs)
{
    g_err_id[0] = g_err_msg[0] = 0;
    g_nblocks = 0;
    g_created = NULL;
    g_auto_freed = 0;
    g_arrays_dropped = 0;
    /* MATLAB always offers plhs[0] (ans), even for nlhs = 0: the caller passes
     * room for max(nlhs, 1) outputs */
    const int nout = nlhs > 0 ? nlhs : 1;
    for (int i = 0; i < nout; ++i) plhs[i] = NULL;
    volatile int raised = 0;
    g_in_call = 1;
    if (setjmp(g_jmp) == 0) {
        mexFunction(nlhs, plhs, nrhs, (const mxArray **)prhs);
    } else {
        raised = 1;
    }
    g_in_call = 0;
    /* MATLAB's clean-up of what the gateway left behind */
    g_auto_freed = g_nblocks;
    for (int i = 0; i < g_nblocks; ++i) free(g_blocks[i]);
    g_nblocks = 0;
    while (g_created) {
        mxArray *a = g_created;
        g_created = a->next_live;
        int out = 0;
        for (int i = 0; i < nout; ++i)
            if (plhs[i] == a) out = 1;
        if (raised || !out) {
            if (!raised) ++g_arrays_dropped;
            for (int i = 0; i < nout; ++i)
                if (plhs[i] == a) plhs[i] = NULL;
            free_array(a);
        }
    }
    return raised;
}

/* mexAtExit handler (clear mex); 1 when one ran */
int shim_exit(void)
{
    if (!g_atexit) return 0;
    void (*fn)(void) = g_atexit;
    g_atexit = NULL;
    fn();
    return 1;
}
