// valG = FSAIP_gen_mex(valA, rPtrA, cPtrA, rPtrG, cPtrG, nfreqs): the MATLAB
// gateway that replaces FSAIP_mexgateway.cpp:8-85 (called at dMesh.m:1384).
//
// FSAI factor values of A_k = Re A + i*k*Im A, k = 0 .. nfreqs-1, for the
// lower triangle tril(A) in CSR (valA, rPtrA, cPtrA) on the pattern (rPtrG,
// cPtrG), 0-based int32 with ascending columns and the diagonal last in
// every row.  nrows = numel(rPtrG) - 1, nnzG = numel(cPtrG); nfreqs is any
// numeric scalar (the reference passes int32).  valG is a host complex
// nfreqs x nnzG array, valG(k+1, p) the value of pattern entry p in G_k, as
// the reference returns it.  Rows of more than 32 pattern entries are an
// error.  Argument-count errors keep the reference's identifiers and
// messages.
#include <cmath>

#include "dogpup_mex_args.hpp"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  using namespace dogpup_mex;
  if (nrhs != 6) mexErrMsgIdAndTxt("CUDA:FSAIP_mex:nrhs", "6 inputs required.");
  if (nlhs != 1) mexErrMsgIdAndTxt("CUDA:FSAIP_mex:nlhs", "1 output required.");
  const char* id = "DOGPUP:FSAIP_gen_mex:arg";
  std::vector<double> at;
  const double* A = complex_doubles(prhs[0], id, "valA", at);
  const int* rA = int32s(prhs[1], id, "rPtrA");
  const int* cA = int32s(prhs[2], id, "cPtrA");
  const int* rG = int32s(prhs[3], id, "rPtrG");
  const int* cG = int32s(prhs[4], id, "cPtrG");
  if (mxGetNumberOfElements(prhs[5]) != 1) mexErrMsgIdAndTxt(id, "nfreqs must be a scalar.");
  const double nfd = mxGetScalar(prhs[5]);
  if (!(nfd >= 1.0 && nfd <= 65536.0 && std::floor(nfd) == nfd))
    mexErrMsgIdAndTxt(id, "nfreqs must be a positive integer (got %g).", nfd);
  const int nfreqs = (int)nfd;
  const int nrg = count(prhs[3], id, "rPtrG");
  if (nrg < 1) mexErrMsgIdAndTxt(id, "rPtrG must have nrows+1 >= 1 elements.");
  const int nrows = nrg - 1;
  if (count(prhs[1], id, "rPtrA") != nrows + 1)
    mexErrMsgIdAndTxt(id, "rPtrA must have numel(rPtrG) = %d elements.", nrows + 1);
  const int nnzA = count(prhs[2], id, "cPtrA");
  if (count(prhs[0], id, "valA") != nnzA)
    mexErrMsgIdAndTxt(id, "valA has %d values for %d column indices.", count(prhs[0], id, "valA"), nnzA);
  const int nnzG = count(prhs[4], id, "cPtrG");
  const mwSize dims[2] = {(mwSize)nfreqs, (mwSize)nnzG};
  plhs[0] = mxCreateNumericArray(2, dims, mxDOUBLE_CLASS, mxCOMPLEX);
  if (nrows == 0 || nnzG == 0) return;
  if (nnzA == 0) mexErrMsgIdAndTxt(id, "tril(A) has no entries but the pattern has %d.", nnzG);
  check_rc(dogpup_fsaip_gen_mex(reinterpret_cast<double*>(mxGetComplexDoubles(plhs[0])), A, rA, cA, rG,
                                cG, nrows, nnzA, nnzG, nfreqs),
           "DOGPUP:FSAIP_gen_mex:lib");
}
