// y = smv_mex(x, f, A, rPtrA, cPtrA): the MATLAB gateway that replaces
// smv_mexgateway.cpp:8-100 (called at dMesh.m:1284,1301,1319).
//
// y(:,c,k) = (Re A + i*f(k)*Im A) x(:,c,k) for the CSR matrix (A, rPtrA,
// cPtrA), 0-based int32 pointers as sparse_csr builds them
// (dMesh.m:1410-1412), and the block x (nodes x columns x frequencies).
// y is a host complex array of x's shape.  Argument-count errors keep the
// reference's identifiers and messages.
#include "dogpup_mex_args.hpp"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  using namespace dogpup_mex;
  if (nrhs != 5) mexErrMsgIdAndTxt("CUDA:fwdSubCSRmex:nrhs", "5 inputs required.");
  if (nlhs != 1) mexErrMsgIdAndTxt("CUDA:fwdSubCSRmex:nlhs", "1 outputs required.");
  const char* id = "DOGPUP:smv_mex:arg";
  const Block b = block_dims(prhs[0], id);
  std::vector<double> xt, at;
  const double* x = complex_doubles(prhs[0], id, "x", xt);
  const double* f = real_doubles(prhs[1], id, "f");
  const double* A = complex_doubles(prhs[2], id, "A", at);
  const int* rp = int32s(prhs[3], id, "rPtrA");
  const int* cp = int32s(prhs[4], id, "cPtrA");
  const int nnz = count(prhs[4], id, "cPtrA");
  if (count(prhs[2], id, "A") != nnz)
    mexErrMsgIdAndTxt(id, "A has %d values for %d column indices.", count(prhs[2], id, "A"), nnz);
  if (count(prhs[3], id, "rPtrA") != b.nrows + 1)
    mexErrMsgIdAndTxt(id, "rPtrA must have size(x,1)+1 = %d elements.", b.nrows + 1);
  if (count(prhs[1], id, "f") < b.nfreqs)
    mexErrMsgIdAndTxt(id, "f must have at least size(x,3) = %d elements.", b.nfreqs);
  plhs[0] = complex_like(prhs[0]);
  if ((size_t)b.nrows * b.ncols * b.nfreqs == 0 || nnz == 0) return;  // empty block or matrix: y = 0
  check_rc(dogpup_smv_mex(reinterpret_cast<double*>(mxGetComplexDoubles(plhs[0])), x, f, A, rp, cp,
                          b.nrows, b.ncols, b.nfreqs, nnz),
           "DOGPUP:smv_mex:lib");
}
