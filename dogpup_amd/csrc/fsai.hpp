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

// Output of the solver-side FSAI kernel: the factor of pattern entry s
// (global, reference order) for frequency k goes to k * nnz + ginv[s] - p0 of
// g, interleaved complex, float when f32 else double.
struct FsaiDirect {
  void* g;
  const int* ginv;
  long long nnz;
  long long p0;
  int f32;
};

// A values: planar (are, aim) when a_il == 0, interleaved in `are` when 1.
// fscale: per-frequency multiplier of Im A, or nullptr for k = 0..nfreqs-1.
// Synchronizes `st` once (row-length class selection).
int fsai_launch(const double* are, const double* aim, int a_il, const int* rptrA,
                const int* cptrA, const int* rptrG, const int* cptrG, int nrows, int nfreqs,
                const double* fscale, FsaiOut out, hipStream_t st);

// Rows of at most 8 entries (rows longer than that are left untouched), A's
// lower-triangle entries on every row's pattern given by pos (36 per row,
// packed (a, b) -> a(a+1)/2 + b, -1 where A has none); planar A values; no
// synchronisation.
int fsai_pos_launch(const double* are, const double* aim, const int* pos, const int* rptrG, int nrows, int nfreqs,
                    const double* fscale, FsaiDirect out, hipStream_t st);

}  // namespace dogpup
