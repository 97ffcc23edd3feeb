// Multi-RHS CSR kernels for the batched solver.
//
// Solver layout: a block of RHS is a set of groups, one per (frequency,
// column block); group g holds NSP columns for every node, element (g, node,
// c) at ((g*nn + node)*NSP + c) complex.  One 16-byte pack holds CPL complex
// values (2 complex64 or 1 complex128); LPR = NSP/CPL lanes cover one node
// row, so the LPR lanes of a row gather one contiguous NSP*sizeof(complex)
// segment of x per matrix entry (256 B for 32 complex64 columns) and every
// matrix entry is read once for the whole group.  A 256-thread block owns a
// tile of TILE consecutive rows of one group; blocks are ordered group-major,
// so the few groups in flight at a time keep their x working set (~70 MB for
// 269k nodes x 32 complex64 columns) in the 256 MB Infinity Cache while the
// 7/15-point stencil re-reads it.
#pragma once
#include "common.hpp"

namespace dogpup {

constexpr int kTile = 512;  // rows per block

template <typename T, int NSP>
struct Geo {
  static constexpr int CPL = Pack<T>::N;
  static constexpr int LPR = NSP / CPL;
  static constexpr int RPW = 64 / LPR;
  static constexpr int RPB = 4 * RPW;
  static constexpr int STEPS = kTile / RPB;
  static_assert(LPR >= 1 && LPR <= 64, "group width");
  static_assert(kTile % RPB == 0, "tile");
};

// complex double partial-sum accumulator for CPL columns
template <int CPL>
struct Acc {
  double re[CPL], im[CPL];
  __device__ inline void zero() {
#pragma unroll
    for (int q = 0; q < CPL; ++q) { re[q] = 0.0; im[q] = 0.0; }
  }
};

// Reduce lane accumulators of one block into part[(gb)*NSP + c] (complex
// double, interleaved).  `sh` holds 4*64 complex doubles.
template <typename T, int NSP>
__device__ inline void block_reduce_store(Acc<Pack<T>::N>& a, double* sh, double* part, long long gb) {
  using G = Geo<T, NSP>;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int rsub = lane / G::LPR, cl = lane % G::LPR;
#pragma unroll
  for (int q = 0; q < G::CPL; ++q) {
#pragma unroll
    for (int o = G::LPR; o < 64; o <<= 1) {
      a.re[q] += __shfl_xor(a.re[q], o);
      a.im[q] += __shfl_xor(a.im[q], o);
    }
  }
  if (rsub == 0) {
#pragma unroll
    for (int q = 0; q < G::CPL; ++q) {
      sh[2 * ((w * NSP) + cl * G::CPL + q)] = a.re[q];
      sh[2 * ((w * NSP) + cl * G::CPL + q) + 1] = a.im[q];
    }
  }
  __syncthreads();
  if (threadIdx.x < NSP) {
    const int c = threadIdx.x;
    double sr = 0.0, si = 0.0;
#pragma unroll
    for (int ww = 0; ww < 4; ++ww) {
      sr += sh[2 * (ww * NSP + c)];
      si += sh[2 * (ww * NSP + c) + 1];
    }
    part[2 * (gb * NSP + c)] = sr;
    part[2 * (gb * NSP + c) + 1] = si;
  }
}

}  // namespace dogpup
