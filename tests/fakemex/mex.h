/* A minimal stand-in for MATLAB's mex.h / matrix.h (interleaved complex API,
 * R2018a), written for this project's tests: the gateways under
 * dogpup_amd/mex/ compile against it unchanged, and tests drive their
 * mexFunction from Python through the fm_* calls below.
 *
 * Only what the gateways use exists: numeric arrays of class double and
 * int32, real or complex, with up to 8 dimensions stored column-major as in
 * MATLAB.  mexErrMsgIdAndTxt does not return (it throws, and fm_invoke
 * catches it and reports the identifier and message), as MATLAB's does not.
 * The API names and argument meaning are MATLAB's public C matrix API. */
#ifndef DOGPUP_TEST_FAKEMEX_H
#define DOGPUP_TEST_FAKEMEX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef size_t mwSize;
typedef size_t mwIndex;
typedef struct mxArray_tag mxArray;
typedef enum { mxUNKNOWN_CLASS = 0, mxDOUBLE_CLASS = 6, mxSINGLE_CLASS = 7, mxINT32_CLASS = 12 } mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;
typedef double mxDouble;
typedef int32_t mxInt32;
typedef struct {
  mxDouble real, imag;
} mxComplexDouble;

mwSize mxGetNumberOfDimensions(const mxArray* a);
const mwSize* mxGetDimensions(const mxArray* a);
size_t mxGetNumberOfElements(const mxArray* a);
size_t mxGetM(const mxArray* a);
size_t mxGetN(const mxArray* a);
mxClassID mxGetClassID(const mxArray* a);
bool mxIsDouble(const mxArray* a);
bool mxIsInt32(const mxArray* a);
bool mxIsComplex(const mxArray* a);
bool mxIsEmpty(const mxArray* a);
mxDouble* mxGetDoubles(const mxArray* a);
mxComplexDouble* mxGetComplexDoubles(const mxArray* a);
mxInt32* mxGetInt32s(const mxArray* a);
double mxGetScalar(const mxArray* a);
mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID classid, mxComplexity flag);
void mxDestroyArray(mxArray* a);
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...);
/* the gateway's entry, with C linkage as MATLAB's mex.h declares it */
void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);

/* ---- harness calls (not MATLAB API) ---- */
/* a new array holding a copy of `data` (numel elements of the class, two
 * doubles per element for complex; data may be NULL for zeros) */
mxArray* fm_create(int classid, int complexity, int ndim, const size_t* dims, const void* data);
/* the array's storage: numel elements, interleaved re/im when complex */
void* fm_data(const mxArray* a);
typedef void (*fm_mexfunction)(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);
/* calls fn; returns 0, or 1 with the error identifier and message written
 * to id / msg (each `cap` bytes) when the gateway raised an error */
int fm_invoke(fm_mexfunction fn, int nlhs, mxArray** plhs, int nrhs, const mxArray** prhs, char* id,
              char* msg, int cap);

#ifdef __cplusplus
}
#endif

#endif
