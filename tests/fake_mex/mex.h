/*
 * mex.h -- a minimal stand-in for the MATLAB C matrix/MEX API, for testing the
 * MEX gateways in gpumatfun_amd/mex/ without MATLAB.  Written from the public
 * function names and semantics of the documented API; only what the gateways
 * use.  mexErrMsgIdAndTxt throws (MATLAB longjmps back to the prompt).
 */
#ifndef GM_FAKE_MEX_H
#define GM_FAKE_MEX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef size_t mwSize;
typedef size_t mwIndex;
typedef struct mxArray_tag mxArray;
typedef enum { mxDOUBLE_CLASS = 6, mxUINT64_CLASS = 13 } mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

size_t mxGetM(const mxArray* a);
size_t mxGetN(const mxArray* a);
double* mxGetPr(const mxArray* a);
double* mxGetPi(const mxArray* a);
void* mxGetData(const mxArray* a);
mwIndex* mxGetIr(const mxArray* a);
mwIndex* mxGetJc(const mxArray* a);
mwSize mxGetNzmax(const mxArray* a);
int mxIsDouble(const mxArray* a);
int mxIsSparse(const mxArray* a);
int mxIsUint64(const mxArray* a);
int mxIsEmpty(const mxArray* a);
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
mxArray* mxCreateDoubleScalar(double v);
mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID cls, mxComplexity c);
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...);

#ifdef __cplusplus
}
#endif

#endif
