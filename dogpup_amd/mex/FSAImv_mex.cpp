// y = FSAImv_mex(x, G, rPtrG, cPtrG): the MATLAB gateway that replaces
// FSAImv_mexgateway.cpp:8-100 (called at dMesh.m:1295-1296,1315-1316,1322).
//
// y(:,c,k) = G_k x(:,c,k) for the CSR pattern (rPtrG, cPtrG), 0-based
// int32, with one column of values per frequency: G is nnzG x nfreqs,
// G(p, k) the value of entry p in G_k (dMesh.m:1384 transposes FSAIP_gen_mex's
// output into this form).  nnzG = numel(cPtrG).  y is a host complex array
// of x's shape.  Argument-count errors keep the reference's identifiers and
// messages.
#include "dogpup_mex_args.hpp"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  using namespace dogpup_mex;
  if (nrhs != 4) mexErrMsgIdAndTxt("CUDA:FSAImv:nrhs", "4 inputs required.");
  if (nlhs != 1) mexErrMsgIdAndTxt("CUDA:FSAImv:nlhs", "1 outputs required.");
  const char* id = "DOGPUP:FSAImv_mex:arg";
  const Block b = block_dims(prhs[0], id);
  std::vector<double> xt, gt;
  const double* x = complex_doubles(prhs[0], id, "x", xt);
  const double* G = complex_doubles(prhs[1], id, "G", gt);
  const int* rp = int32s(prhs[2], id, "rPtrG");
  const int* cp = int32s(prhs[3], id, "cPtrG");
  const int nnz = count(prhs[3], id, "cPtrG");
  if ((size_t)count(prhs[1], id, "G") != (size_t)nnz * b.nfreqs)
    mexErrMsgIdAndTxt(id, "G must be numel(cPtrG) x size(x,3) = %d x %d.", nnz, b.nfreqs);
  if (count(prhs[2], id, "rPtrG") != b.nrows + 1)
    mexErrMsgIdAndTxt(id, "rPtrG must have size(x,1)+1 = %d elements.", b.nrows + 1);
  plhs[0] = complex_like(prhs[0]);
  if ((size_t)b.nrows * b.ncols * b.nfreqs == 0 || nnz == 0) return;  // empty block or factor: y = 0
  check_rc(dogpup_fsaimv_mex(reinterpret_cast<double*>(mxGetComplexDoubles(plhs[0])), x, G, rp, cp,
                             b.nrows, b.ncols, b.nfreqs, nnz),
           "DOGPUP:FSAImv_mex:lib");
}
