/*
 * mex.h -- test double of the MEX API subset that festung_amd/mex/festung_mex.cpp
 * uses, so that the gateway compiles and runs in the test suite without
 * Matlab or Octave (neither is installed here).  Written for this project
 * from the MEX API's documented names and signatures; the behaviour is that
 * of tests/mex/mex_standin.cpp (mexErrMsgIdAndTxt throws instead of
 * unwinding into the interpreter).  Not used by any build of the product.
 */
#ifndef FESTUNG_TEST_MEX_H
#define FESTUNG_TEST_MEX_H
#include <stddef.h>
#include <stdbool.h>

typedef size_t mwSize;
typedef struct mxArray_tag mxArray;
typedef bool mxLogical;
typedef enum { mxUNKNOWN_CLASS = 0, mxLOGICAL_CLASS = 3, mxCHAR_CLASS = 4, mxDOUBLE_CLASS = 6,
               mxUINT64_CLASS = 13 } mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

#ifdef __cplusplus
extern "C" {
#endif
bool mxIsChar(const mxArray* a);
bool mxIsDouble(const mxArray* a);
bool mxIsLogical(const mxArray* a);
bool mxIsUint64(const mxArray* a);
bool mxIsComplex(const mxArray* a);
size_t mxGetM(const mxArray* a);
size_t mxGetN(const mxArray* a);
size_t mxGetNumberOfElements(const mxArray* a);
double* mxGetPr(const mxArray* a);
void* mxGetData(const mxArray* a);
mxLogical* mxGetLogicals(const mxArray* a);
double mxGetScalar(const mxArray* a);
int mxGetString(const mxArray* a, char* buf, mwSize buflen);
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
mxArray* mxCreateDoubleScalar(double v);
mxArray* mxCreateLogicalMatrix(mwSize m, mwSize n);
mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity c);
mxArray* mxCreateString(const char* s);
void mxDestroyArray(mxArray* a);
void mexLock(void);
void mexUnlock(void);
int mexAtExit(void (*fn)(void));
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...);
#ifdef __cplusplus
}
void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);
#endif
#endif
