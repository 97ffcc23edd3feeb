#pragma once
#include "common.hpp"

namespace dogpup {

enum { FSAI_OUT_REF = 0, FSAI_OUT_SOLVER = 1 };

// Output of the FSAI kernel.  FSAI_OUT_REF: planar re/im, out[k + nfreqs*j]
// (the reference launcher's layout).  FSAI_OUT_SOLVER: interleaved complex in
// re, [k][nnzG].
struct FsaiOut {
  double* re;
  double* im;
  long long nnzG;
  int layout;
};

// A values: planar (are, aim) when a_il == 0, interleaved in `are` when 1.
// fscale: per-frequency multiplier of Im A, or nullptr for k = 0..nfreqs-1.
// Synchronizes `st` once (row-length class selection).
int fsai_launch(const double* are, const double* aim, int a_il, const int* rptrA,
                const int* cptrA, const int* rptrG, const int* cptrG, int nrows, int nfreqs,
                const double* fscale, FsaiOut out, hipStream_t st);

}  // namespace dogpup
