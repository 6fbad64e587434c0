/*
 * matrix.h -- a minimal stand-in for MATLAB's mx array API, written for the
 * tests of the MEX gateways in slim-release-apps_amd/mex/.  It declares only
 * the public names those gateways use (double matrices with separate real and
 * imaginary storage, the pre-R2018a layout the reference MEX files assume) and
 * is implemented in mexshim.c.  Test infrastructure: never part of a MATLAB
 * build, which uses MATLAB's own headers.
 */
#ifndef MEXSHIM_MATRIX_H
#define MEXSHIM_MATRIX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef size_t mwSize;
typedef size_t mwIndex;
typedef struct mxArray_tag mxArray;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

double *mxGetPr(const mxArray *a);
double *mxGetPi(const mxArray *a);
size_t mxGetM(const mxArray *a);
size_t mxGetN(const mxArray *a);
size_t mxGetNumberOfElements(const mxArray *a);
int mxIsComplex(const mxArray *a);
int mxIsDouble(const mxArray *a);
double mxGetScalar(const mxArray *a);
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
void *mxCalloc(size_t n, size_t size);
void mxFree(void *p);
void mxDestroyArray(mxArray *a);

#ifdef __cplusplus
}
#endif
#endif
