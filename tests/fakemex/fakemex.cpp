// The array store and error path behind tests/fakemex/mex.h.
#include "mex.h"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

struct mxArray_tag {
  mxClassID cls;
  bool cplx;
  std::vector<mwSize> dims;
  std::vector<unsigned char> bytes;
};

namespace {

struct MexError {
  std::string id, msg;
};

size_t elem_bytes(const mxArray* a) {
  const size_t b = a->cls == mxDOUBLE_CLASS ? 8 : a->cls == mxSINGLE_CLASS ? 4 : 4;
  return a->cplx ? 2 * b : b;
}

}  // namespace

extern "C" {

mwSize mxGetNumberOfDimensions(const mxArray* a) { return a->dims.size(); }
const mwSize* mxGetDimensions(const mxArray* a) { return a->dims.data(); }
size_t mxGetNumberOfElements(const mxArray* a) {
  size_t n = 1;
  for (mwSize d : a->dims) n *= d;
  return n;
}
size_t mxGetM(const mxArray* a) { return a->dims[0]; }
size_t mxGetN(const mxArray* a) { return mxGetNumberOfElements(a) / (a->dims[0] ? a->dims[0] : 1); }
mxClassID mxGetClassID(const mxArray* a) { return a->cls; }
bool mxIsDouble(const mxArray* a) { return a->cls == mxDOUBLE_CLASS; }
bool mxIsInt32(const mxArray* a) { return a->cls == mxINT32_CLASS; }
bool mxIsComplex(const mxArray* a) { return a->cplx; }
bool mxIsEmpty(const mxArray* a) { return mxGetNumberOfElements(a) == 0; }

// MATLAB returns NULL when the class or complexity does not match, and for
// empty arrays
mxDouble* mxGetDoubles(const mxArray* a) {
  if (a->cls != mxDOUBLE_CLASS || a->cplx || a->bytes.empty()) return nullptr;
  return (mxDouble*)a->bytes.data();
}
mxComplexDouble* mxGetComplexDoubles(const mxArray* a) {
  if (a->cls != mxDOUBLE_CLASS || !a->cplx || a->bytes.empty()) return nullptr;
  return (mxComplexDouble*)a->bytes.data();
}
mxInt32* mxGetInt32s(const mxArray* a) {
  if (a->cls != mxINT32_CLASS || a->cplx || a->bytes.empty()) return nullptr;
  return (mxInt32*)a->bytes.data();
}

// the real part of the first element, converted to double; an empty array
// is an error in MATLAB
double mxGetScalar(const mxArray* a) {
  if (a->bytes.empty()) throw MexError{"MATLAB:mxGetScalar:empty", "mxGetScalar of an empty array"};
  if (a->cls == mxDOUBLE_CLASS) return *(const double*)a->bytes.data();
  if (a->cls == mxINT32_CLASS) return (double)*(const int32_t*)a->bytes.data();
  return (double)*(const float*)a->bytes.data();
}

mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID classid, mxComplexity flag) {
  mxArray* a = new mxArray_tag{classid, flag == mxCOMPLEX, {}, {}};
  a->dims.assign(dims, dims + ndim);
  while (a->dims.size() > 2 && a->dims.back() == 1) a->dims.pop_back();  // trailing singletons drop
  while (a->dims.size() < 2) a->dims.push_back(1);
  a->bytes.assign(mxGetNumberOfElements(a) * elem_bytes(a), 0);
  return a;
}

void mxDestroyArray(mxArray* a) { delete a; }

void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) {
  char buf[2048];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw MexError{id, buf};
}

mxArray* fm_create(int classid, int complexity, int ndim, const size_t* dims, const void* data) {
  mxArray* a = mxCreateNumericArray((mwSize)ndim, dims, (mxClassID)classid, (mxComplexity)complexity);
  if (data && !a->bytes.empty()) std::memcpy(a->bytes.data(), data, a->bytes.size());
  return a;
}

void* fm_data(const mxArray* a) { return a->bytes.empty() ? nullptr : (void*)a->bytes.data(); }

int fm_invoke(fm_mexfunction fn, int nlhs, mxArray** plhs, int nrhs, const mxArray** prhs, char* id,
              char* msg, int cap) {
  try {
    fn(nlhs, plhs, nrhs, prhs);
    return 0;
  } catch (const MexError& e) {
    snprintf(id, cap, "%s", e.id.c_str());
    snprintf(msg, cap, "%s", e.msg.c_str());
    return 1;
  }
}

}  // extern "C"
