// Argument handling shared by the three MATLAB gateways (smv_mex.cpp,
// FSAImv_mex.cpp, FSAIP_gen_mex.cpp).  Build each gateway with MATLAB's
// interleaved complex API:
//
//   mex -R2018a -I<repo>/include smv_mex.cpp -L<repo>/dogpup_amd/lib -ldogpup_hip
//
// The reference's gateways take gpuArrays (smv_mexgateway.cpp:22-39); these
// take host arrays and hand them to the mex-level entry points of
// libdogpup_hip, which stage them on the GPU.  Where the reference reads an
// argument without checking it, these check it and raise a MATLAB error:
//  * a block x of 2 dimensions is one frequency (the reference reads
//    dims[2] past the end of a 2-element dimension array);
//  * counts of non-zeros come from numel(cPtr), whatever its orientation
//    (FSAImv_mexgateway.cpp:44 takes dimsC[0], the row count);
//  * a real array stands for a complex one with zero imaginary parts, as
//    mxGPUCopyImag of a real gpuArray gives zeros;
//  * index arrays must be int32 and the CSR structure consistent (the
//    library checks it before launching anything).
#ifndef DOGPUP_MEX_ARGS_HPP
#define DOGPUP_MEX_ARGS_HPP

#include <vector>

#include "dogpup.h"
#include "mex.h"

namespace dogpup_mex {

// Interleaved complex doubles of a double array.  A real array is widened
// into `tmp` with zero imaginary parts.
inline const double* complex_doubles(const mxArray* a, const char* id, const char* name,
                                     std::vector<double>& tmp) {
  if (!mxIsDouble(a)) mexErrMsgIdAndTxt(id, "%s must be of class double.", name);
  const size_t n = mxGetNumberOfElements(a);
  if (mxIsComplex(a)) return reinterpret_cast<const double*>(mxGetComplexDoubles(a));
  tmp.assign(2 * n, 0.0);
  const double* r = mxGetDoubles(a);
  for (size_t i = 0; i < n; ++i) tmp[2 * i] = r[i];
  return tmp.data();
}

inline const double* real_doubles(const mxArray* a, const char* id, const char* name) {
  if (!mxIsDouble(a) || mxIsComplex(a)) mexErrMsgIdAndTxt(id, "%s must be a real double array.", name);
  return mxGetDoubles(a);
}

inline const int* int32s(const mxArray* a, const char* id, const char* name) {
  if (!mxIsInt32(a) || mxIsComplex(a)) mexErrMsgIdAndTxt(id, "%s must be a real int32 array.", name);
  return reinterpret_cast<const int*>(mxGetInt32s(a));
}

inline int count(const mxArray* a, const char* id, const char* name) {
  const size_t n = mxGetNumberOfElements(a);
  if (n > 0x7fffffffu) mexErrMsgIdAndTxt(id, "%s has more than 2^31-1 elements.", name);
  return (int)n;
}

// nrows x ncols x nfreqs of a block x (reference layout, x(row, col, freq));
// a 2-D x is one frequency
struct Block {
  int nrows, ncols, nfreqs;
};

inline Block block_dims(const mxArray* x, const char* id) {
  const mwSize nd = mxGetNumberOfDimensions(x);
  const mwSize* d = mxGetDimensions(x);
  if (nd > 3) mexErrMsgIdAndTxt(id, "x must have at most 3 dimensions (nodes x columns x frequencies).");
  if (d[0] > 0x7fffffffu || d[1] > 0x7fffffffu || (nd > 2 && d[2] > 0x7fffffffu))
    mexErrMsgIdAndTxt(id, "x has a dimension above 2^31-1.");
  return Block{(int)d[0], (int)d[1], nd > 2 ? (int)d[2] : 1};
}

// a complex double array of x's shape for the result
inline mxArray* complex_like(const mxArray* x) {
  return mxCreateNumericArray(mxGetNumberOfDimensions(x), mxGetDimensions(x), mxDOUBLE_CLASS, mxCOMPLEX);
}

inline void check_rc(int rc, const char* id) {
  if (rc != DOGPUP_OK) mexErrMsgIdAndTxt(id, "%s", dogpup_last_error());
}

}  // namespace dogpup_mex

#endif
