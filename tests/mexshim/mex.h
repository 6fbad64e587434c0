/*
 * mex.h -- minimal stand-in for MATLAB's MEX entry API (see matrix.h): the
 * gateway signature, error raising through a long jump back to the caller of
 * mexFunction, and exit handlers.  Implemented in mexshim.c; test
 * infrastructure only.
 */
#ifndef MEXSHIM_MEX_H
#define MEXSHIM_MEX_H

#include "matrix.h"

#ifdef __cplusplus
extern "C" {
#endif

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);
void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...);
int mexAtExit(void (*fn)(void));
int mexPrintf(const char *fmt, ...);

#ifdef __cplusplus
}
#endif
#endif
