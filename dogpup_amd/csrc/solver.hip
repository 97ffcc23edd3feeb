// Batched FSAI-preconditioned BiCGSTAB over every (frequency, column) RHS.
//
// Algorithm (behaviour of DOGPUP/dMesh.m:1247-1348, gpuBicstab_FSAIP):
// right-preconditioned BiCGSTAB on A_k x = b with M^-1 = G_k^T G_k, x0 = 0,
// r_hat0 = r0, unconjugated bilinear forms, omega = (Gt . Gs)/(Gt . Gt)
// weighted by G, convergence on ||s||/||b|| or ||r||/||b||.
//
// Three forms share the CSR kernels of kernels.hpp:
//  * DOGPUP_MODE_C128: complex128, fused passes (dots computed inside the
//    SpMV/update that produces the operand), per-RHS convergence with
//    groups skipped once all their columns have converged.
//  * DOGPUP_MODE_C64_REFINE: the same fused loop in complex64 as the inner
//    solver of FP64 iterative refinement: r64 = b - A x64 in FP64, inner
//    solve A d = r64 in complex64 to a relative tolerance, x64 += d.
//  * DOGPUP_MODE_C128_PLAIN: complex128 with the reference's structure:
//    separate G, G^T and A passes, separate dot/axpy passes with
//    temporaries, the reciprocal step `alph`, lockstep convergence on the
//    maximum relative residual checked by the host twice per iteration.
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <algorithm>
#include <map>
#include <memory>
#include <vector>

#include "fsai.hpp"
#include "kernels.hpp"

namespace dogpup {

enum { MAT_A = 0, MAT_G = 1 };
// SpMV epilogues: EPI_STORE y; EPI_STORE_DOT y and w.y; EPI_DOT2 only w.y
// and y.y (y not stored); EPI_STORE_SELF y and y.y; EPI_AXPBY y <- y_new +
// coef*y_old (in place, one complex coefficient per column); EPI_AXPBY_X
// also xacc += coef2*y_old.  All dots are unconjugated.
enum { EPI_STORE = 0, EPI_STORE_DOT = 1, EPI_DOT2 = 2, EPI_STORE_SELF = 3, EPI_AXPBY = 4, EPI_AXPBY_X = 5 };
enum { ST_ACTIVE = -1, ST_SKIP = -2 };

// Row tiles are assigned so that each XCD sweeps one contiguous range of
// tiles: workgroups are dispatched round-robin over the 8 XCDs, and
// gridDim.x is a multiple of 8, so blockIdx.x % 8 names the XCD.  A tile's
// stencil neighbours then sit in the same XCD's L2.
__device__ inline int xcd_tile(int bx) { return (bx & 7) * (gridDim.x >> 3) + (bx >> 3); }

// ------------------------------------------------------------------ SpMV ---
// y = M x for every group; MAT_A: M = v0 + i*fscale[f]*v1 (planar, shared by
// all frequencies); MAT_G: M = per-frequency complex values v0[f][nnz].
//
// The block walks its row tile in chunks of kChunkRows rows.  Each chunk's
// CSR row pointers, column indices and values are first copied into LDS with
// coalesced loads (fetching the matrix entry per wavefront from global
// memory would cost as many address cycles as the x gather itself); the
// only vector-memory instruction per nonzero is then the 16-byte-per-lane
// gather of x.  A row issues U gathers before it consumes any (U = 8 by
// default: 6 waves per SIMD with 8 loads in flight each measured faster than
// 4 or 16 on the 15-point A and 8-point G rows); the entries past the row's
// end gather the row's last column again with weight zero, which keeps the
// loads unconditional.  A chunk with more
// than kChunkCap nonzeros reads the CSR arrays from global memory instead.
// Per-lane column coefficients (complex double per RHS) as T.
template <typename T, int CPL>
__device__ inline void load_coef(const double* __restrict__ c, long long rhs0, T re[CPL], T im[CPL]) {
#pragma unroll
  for (int e = 0; e < CPL; ++e) {
    re[e] = (T)c[2 * (rhs0 + e)];
    im[e] = (T)c[2 * (rhs0 + e) + 1];
  }
}

constexpr int kChunkRows = 64;
constexpr int kChunkCap = 1024;

// Row products of one row tile [tile0, tile1): for every row, (ar, ai) =
// (M x)(row) over this lane's CPL columns, handed to epi(row, ar, ai).
// srp/scol/sval are the caller's LDS chunk buffers.  All threads of the block
// call it (it synchronises).
template <typename T, int NSP, int MAT, int U, typename Epi>
__device__ inline void spmv_rows(const int* __restrict__ rptr, const int* __restrict__ cptr,
                                 const T* __restrict__ v0, const T* __restrict__ v1, T fk,
                                 const T* __restrict__ xg, int tile0, int tile1, int* srp, int* scol, T* sval,
                                 Epi&& epi) {
  using G = Geo<T, NSP>;
  constexpr int CPL = G::CPL;
  static_assert(kChunkRows % G::RPB == 0, "chunk");
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int rsub = lane / G::LPR;
  for (int r0 = tile0; r0 < tile1; r0 += kChunkRows) {
    const int r1 = min(r0 + kChunkRows, tile1);
    __syncthreads();  // the previous chunk's LDS reads are done
    for (int q = threadIdx.x; q <= r1 - r0; q += 256) srp[q] = rptr[r0 + q];
    __syncthreads();
    const int es = srp[0], ne = srp[r1 - r0] - es;
    const bool staged = ne <= kChunkCap;
    if (staged) {
      for (int q = threadIdx.x; q < ne; q += 256) {
        scol[q] = cptr[es + q];
        if (MAT == MAT_A) {
          sval[2 * q] = v0[es + q];
          sval[2 * q + 1] = fk * v1[es + q];
        } else {
          sval[2 * q] = v0[2 * (long long)(es + q)];
          sval[2 * q + 1] = v0[2 * (long long)(es + q) + 1];
        }
      }
    }
    __syncthreads();
#pragma unroll 1
    for (int st = 0; st < kChunkRows / G::RPB; ++st) {
      const int lr = st * G::RPB + wv * G::RPW + rsub;
      const int row = r0 + lr;
      if (row >= r1) continue;
      T ar[CPL], ai[CPL];
#pragma unroll
      for (int e = 0; e < CPL; ++e) { ar[e] = T(0); ai[e] = T(0); }
      const int s0 = srp[lr] - es, s1 = srp[lr + 1] - es;
      for (int p = s0; p < s1; p += U) {
        int cc[U];
        T mr[U], mi[U];
#pragma unroll
        for (int q = 0; q < U; ++q) {
          const int pq = min(p + q, s1 - 1);
          if (staged) {
            cc[q] = scol[pq];
            mr[q] = sval[2 * pq];
            mi[q] = sval[2 * pq + 1];
          } else {
            const long long pg = es + pq;
            cc[q] = cptr[pg];
            if (MAT == MAT_A) { mr[q] = v0[pg]; mi[q] = fk * v1[pg]; }
            else { mr[q] = v0[2 * pg]; mi[q] = v0[2 * pg + 1]; }
          }
          if (p + q >= s1) { mr[q] = T(0); mi[q] = T(0); }
        }
        T xr[U][CPL], xi[U][CPL];
#pragma unroll
        for (int q = 0; q < U; ++q) Pack<T>::load(xg + (size_t)cc[q] * NSP * 2, xr[q], xi[q]);
#pragma unroll
        for (int q = 0; q < U; ++q)
#pragma unroll
          for (int e = 0; e < CPL; ++e) {
            ar[e] += mr[q] * xr[q][e] - mi[q] * xi[q][e];
            ai[e] += mr[q] * xi[q][e] + mi[q] * xr[q][e];
          }
      }
      epi(row, ar, ai);
    }
  }
}

// 6 waves per SIMD (at most 80 VGPRs) measured best for U = 8.
template <typename T, int NSP, int MAT, int EPI, int U>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(U <= 8 ? 6 : 4))) void k_spmv(
    const int* __restrict__ rptr, const int* __restrict__ cptr, const T* __restrict__ v0,
    const T* __restrict__ v1, long long nnz, const double* __restrict__ fscale,
    const T* __restrict__ x, T* __restrict__ y, const T* __restrict__ w, const double* __restrict__ coef,
    const double* __restrict__ coef2, T* __restrict__ xacc, double* __restrict__ part1,
    double* __restrict__ part2, int nn, int ncb, int g0, int nbx, int tile_rows,
    const int* __restrict__ grp_active) {
  using G = Geo<T, NSP>;
  constexpr int CPL = G::CPL;
  const int g = blockIdx.y, bx = blockIdx.x;
  if (grp_active && !grp_active[g]) return;
  __shared__ double sh[2 * 4 * NSP];
  __shared__ int srp[kChunkRows + 1];
  __shared__ int scol[kChunkCap];
  __shared__ T sval[2 * kChunkCap];
  const int lane = threadIdx.x & 63;
  const int c0 = (lane % G::LPR) * CPL;
  const int f = (g0 + g) / ncb;
  T fk = T(0);
  const T* gv = v0;
  if (MAT == MAT_A) fk = (T)fscale[f];
  else gv = v0 + 2 * (long long)f * nnz;
  const size_t gbase = (size_t)g * nn;
  Acc<CPL> d1, d2;
  d1.zero();
  d2.zero();
  constexpr bool AXPBY = EPI == EPI_AXPBY || EPI == EPI_AXPBY_X;
  T cr[CPL], ci[CPL], c2r[CPL], c2i[CPL];
  if (AXPBY) load_coef<T, CPL>(coef, (long long)g * NSP + c0, cr, ci);
  if (EPI == EPI_AXPBY_X) load_coef<T, CPL>(coef2, (long long)g * NSP + c0, c2r, c2i);
  const int tile0 = xcd_tile(bx) * tile_rows;
  const int tile1 = min(tile0 + tile_rows, nn);
  spmv_rows<T, NSP, MAT, U>(rptr, cptr, gv, v1, fk, x + (gbase * NSP + c0) * 2, tile0, tile1, srp, scol, sval,
                            [&](int row, T* ar, T* ai) {
    const size_t off = ((gbase + row) * NSP + c0) * 2;
    if (AXPBY) {
      T pr[CPL], pi[CPL];
      Pack<T>::load(y + off, pr, pi);
      if (EPI == EPI_AXPBY_X) {
        T xr[CPL], xi[CPL];
        Pack<T>::load(xacc + off, xr, xi);
#pragma unroll
        for (int e = 0; e < CPL; ++e) {
          xr[e] += c2r[e] * pr[e] - c2i[e] * pi[e];
          xi[e] += c2r[e] * pi[e] + c2i[e] * pr[e];
        }
        Pack<T>::store(xacc + off, xr, xi);
      }
#pragma unroll
      for (int e = 0; e < CPL; ++e) {
        ar[e] += cr[e] * pr[e] - ci[e] * pi[e];
        ai[e] += cr[e] * pi[e] + ci[e] * pr[e];
      }
    }
    if (EPI != EPI_DOT2) Pack<T>::store(y + off, ar, ai);
    if (EPI == EPI_STORE_SELF) {
#pragma unroll
      for (int e = 0; e < CPL; ++e) {
        d1.re[e] += (double)ar[e] * ar[e] - (double)ai[e] * ai[e];
        d1.im[e] += 2.0 * (double)ar[e] * ai[e];
      }
    }
    if (EPI == EPI_STORE_DOT || EPI == EPI_DOT2) {
      T wr[CPL], wi[CPL];
      Pack<T>::load(w + off, wr, wi);
#pragma unroll
      for (int e = 0; e < CPL; ++e) {
        d1.re[e] += (double)wr[e] * ar[e] - (double)wi[e] * ai[e];
        d1.im[e] += (double)wr[e] * ai[e] + (double)wi[e] * ar[e];
        if (EPI == EPI_DOT2) {
          d2.re[e] += (double)ar[e] * ar[e] - (double)ai[e] * ai[e];
          d2.im[e] += 2.0 * (double)ar[e] * ai[e];
        }
      }
    }
  });
  if (EPI != EPI_STORE && !AXPBY) {
    block_reduce_store<T, NSP>(d1, sh, part1, (long long)g * nbx + bx);
    if (EPI == EPI_DOT2) {
      __syncthreads();
      block_reduce_store<T, NSP>(d2, sh, part2, (long long)g * nbx + bx);
    }
  }
}

#define ROW_LOOP_BEGIN                                                         \
  using G = Geo<T, NSP>;                                                       \
  constexpr int CPL = G::CPL;                                                  \
  const int g = blockIdx.y, bx = blockIdx.x;                                   \
  if (grp_active && !grp_active[g]) return;                                    \
  __shared__ double sh[2 * 4 * NSP];                                           \
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;                    \
  const int rsub = lane / G::LPR, cl = lane % G::LPR, c0 = cl * CPL;           \
  const long long rhs0 = (long long)g * NSP + c0;                              \
  const size_t gbase = (size_t)g * nn;                                         \
  (void)sh; (void)rhs0;                                                        \
  for (int st = 0; st < tile_rows / G::RPB; ++st) {                            \
    const int row = xcd_tile(bx) * tile_rows + st * G::RPB + wv * G::RPW + rsub; \
    if (row >= nn) break;                                                      \
    const size_t off = ((gbase + row) * NSP + c0) * 2;

#define ROW_LOOP_END }

// s = r - alpha v ; part1 = sum |s|^2
template <typename T, int NSP>
__global__ __launch_bounds__(256) void k_axpy_s(const T* __restrict__ r, const T* __restrict__ v,
                                                T* __restrict__ s, const double* __restrict__ alpha,
                                                double* __restrict__ part1, int nn, int nbx, int tile_rows,
                                                const int* __restrict__ grp_active) {
  Acc<Geo<T, NSP>::CPL> d1;
  d1.zero();
  T alr[Geo<T, NSP>::CPL], ali[Geo<T, NSP>::CPL];
  ROW_LOOP_BEGIN
  if (st == 0) load_coef<T, CPL>(alpha, rhs0, alr, ali);
  T rr[CPL], ri[CPL], vr[CPL], vi[CPL];
  Pack<T>::load(r + off, rr, ri);
  Pack<T>::load(v + off, vr, vi);
#pragma unroll
  for (int e = 0; e < CPL; ++e) {
    rr[e] -= alr[e] * vr[e] - ali[e] * vi[e];
    ri[e] -= alr[e] * vi[e] + ali[e] * vr[e];
    d1.re[e] += (double)rr[e] * rr[e] + (double)ri[e] * ri[e];
  }
  Pack<T>::store(s + off, rr, ri);
  ROW_LOOP_END
  block_reduce_store<T, NSP>(d1, sh, part1, (long long)g * nbx + bx);
}

// x += alpha h + omega z (active columns only); r = s - omega t;
// part1 = sum rhat0 r, part2 = sum |r|^2
template <typename T, int NSP>
__global__ __launch_bounds__(256) void k_update_xr(
    T* __restrict__ x, const T* __restrict__ h, const T* __restrict__ z, T* __restrict__ r,
    const T* __restrict__ s, const T* __restrict__ t, const T* __restrict__ rhat,
    const double* __restrict__ alpha, const double* __restrict__ omega,
    const int* __restrict__ status, double* __restrict__ part1, double* __restrict__ part2, int nn,
    int nbx, int tile_rows, const int* __restrict__ grp_active) {
  Acc<Geo<T, NSP>::CPL> d1, d2;
  d1.zero();
  d2.zero();
  T alr[Geo<T, NSP>::CPL], ali[Geo<T, NSP>::CPL], omr[Geo<T, NSP>::CPL], omi[Geo<T, NSP>::CPL];
  bool act[Geo<T, NSP>::CPL];
  ROW_LOOP_BEGIN
  if (st == 0) {
    load_coef<T, CPL>(alpha, rhs0, alr, ali);
    load_coef<T, CPL>(omega, rhs0, omr, omi);
#pragma unroll
    for (int e = 0; e < CPL; ++e) act[e] = status[rhs0 + e] == ST_ACTIVE;
  }
  T xr[CPL], xi[CPL], hr[CPL], hi[CPL], zr[CPL], zi[CPL];
  T sr[CPL], si[CPL], tr[CPL], ti[CPL], qr[CPL], qi[CPL];
  Pack<T>::load(x + off, xr, xi);
  Pack<T>::load(h + off, hr, hi);
  Pack<T>::load(z + off, zr, zi);
  Pack<T>::load(s + off, sr, si);
  Pack<T>::load(t + off, tr, ti);
  Pack<T>::load(rhat + off, qr, qi);
#pragma unroll
  for (int e = 0; e < CPL; ++e) {
    if (act[e]) {
      xr[e] += alr[e] * hr[e] - ali[e] * hi[e] + omr[e] * zr[e] - omi[e] * zi[e];
      xi[e] += alr[e] * hi[e] + ali[e] * hr[e] + omr[e] * zi[e] + omi[e] * zr[e];
    }
    sr[e] -= omr[e] * tr[e] - omi[e] * ti[e];
    si[e] -= omr[e] * ti[e] + omi[e] * tr[e];
    d1.re[e] += (double)qr[e] * sr[e] - (double)qi[e] * si[e];
    d1.im[e] += (double)qr[e] * si[e] + (double)qi[e] * sr[e];
    d2.re[e] += (double)sr[e] * sr[e] + (double)si[e] * si[e];
  }
  Pack<T>::store(x + off, xr, xi);
  Pack<T>::store(r + off, sr, si);
  ROW_LOOP_END
  block_reduce_store<T, NSP>(d1, sh, part1, (long long)g * nbx + bx);
  __syncthreads();
  block_reduce_store<T, NSP>(d2, sh, part2, (long long)g * nbx + bx);
}

// p = r + beta (p - omega v)
template <typename T, int NSP>
__global__ __launch_bounds__(256) void k_update_p(T* __restrict__ p, const T* __restrict__ r,
                                                  const T* __restrict__ v,
                                                  const double* __restrict__ beta,
                                                  const double* __restrict__ omega, int nn, int nbx, int tile_rows,
                                                  const int* __restrict__ grp_active) {
  T br[Geo<T, NSP>::CPL], bi[Geo<T, NSP>::CPL], omr[Geo<T, NSP>::CPL], omi[Geo<T, NSP>::CPL];
  ROW_LOOP_BEGIN
  if (st == 0) {
    load_coef<T, CPL>(beta, rhs0, br, bi);
    load_coef<T, CPL>(omega, rhs0, omr, omi);
  }
  T pr[CPL], pi[CPL], rr[CPL], ri[CPL], vr[CPL], vi[CPL];
  Pack<T>::load(p + off, pr, pi);
  Pack<T>::load(r + off, rr, ri);
  Pack<T>::load(v + off, vr, vi);
#pragma unroll
  for (int e = 0; e < CPL; ++e) {
    const T ur = pr[e] - (omr[e] * vr[e] - omi[e] * vi[e]);
    const T ui = pi[e] - (omr[e] * vi[e] + omi[e] * vr[e]);
    pr[e] = rr[e] + br[e] * ur - bi[e] * ui;
    pi[e] = ri[e] + br[e] * ui + bi[e] * ur;
  }
  Pack<T>::store(p + off, pr, pi);
  ROW_LOOP_END
  (void)bx;
}

// COCG residual step: r -= alpha q; part2 = sum |r|^2.  (x += alpha p is
// applied in the G^T pass of the same iteration, which reads p anyway.)
template <typename T, int NSP>
__global__ __launch_bounds__(256) void k_cg_update(T* __restrict__ r, const T* __restrict__ q,
                                                   const double* __restrict__ alpha, double* __restrict__ part2,
                                                   int nn, int nbx, int tile_rows,
                                                   const int* __restrict__ grp_active) {
  Acc<Geo<T, NSP>::CPL> d2;
  d2.zero();
  T alr[Geo<T, NSP>::CPL], ali[Geo<T, NSP>::CPL];
  ROW_LOOP_BEGIN
  if (st == 0) load_coef<T, CPL>(alpha, rhs0, alr, ali);
  T rr[CPL], ri[CPL], qr[CPL], qi[CPL];
  Pack<T>::load(r + off, rr, ri);
  Pack<T>::load(q + off, qr, qi);
#pragma unroll
  for (int e = 0; e < CPL; ++e) {
    rr[e] -= alr[e] * qr[e] - ali[e] * qi[e];
    ri[e] -= alr[e] * qi[e] + ali[e] * qr[e];
    d2.re[e] += (double)rr[e] * rr[e] + (double)ri[e] * ri[e];
  }
  Pack<T>::store(r + off, rr, ri);
  ROW_LOOP_END
  block_reduce_store<T, NSP>(d2, sh, part2, (long long)g * nbx + bx);
}

// r = b, rhat = b, p = b, x = 0; part1 = sum |b|^2, part2 = sum b b
template <typename T, int NSP>
__global__ __launch_bounds__(256) void k_init(const T* __restrict__ b, T* __restrict__ x,
                                              T* __restrict__ r, T* __restrict__ rhat,
                                              T* __restrict__ p, double* __restrict__ part1,
                                              double* __restrict__ part2, int nn, int nbx, int tile_rows,
                                              const int* __restrict__ grp_active) {
  Acc<Geo<T, NSP>::CPL> d1, d2;
  d1.zero();
  d2.zero();
  ROW_LOOP_BEGIN
  T br[CPL], bi[CPL], zr[CPL], zi[CPL];
  Pack<T>::load(b + off, br, bi);
#pragma unroll
  for (int e = 0; e < CPL; ++e) {
    zr[e] = T(0);
    zi[e] = T(0);
    d1.re[e] += (double)br[e] * br[e] + (double)bi[e] * bi[e];
    d2.re[e] += (double)br[e] * br[e] - (double)bi[e] * bi[e];
    d2.im[e] += 2.0 * (double)br[e] * bi[e];
  }
  Pack<T>::store(x + off, zr, zi);
  if (r != b) Pack<T>::store(r + off, br, bi);
  Pack<T>::store(rhat + off, br, bi);
  Pack<T>::store(p + off, br, bi);
  ROW_LOOP_END
  block_reduce_store<T, NSP>(d1, sh, part1, (long long)g * nbx + bx);
  __syncthreads();
  block_reduce_store<T, NSP>(d2, sh, part2, (long long)g * nbx + bx);
}

// FP64 refinement residual: r64 = b - A x64 (complex128), with A x64 from
// the staged row products of k_spmv (x64 == nullptr: x64 = 0, r64 = b).
// When `write`, stores complex64 r, p (and rhat, if given) = r64 and x32 = 0
// for columns whose outer status is active (zeros elsewhere);
// part1 = sum |r64|^2 (FP64), part2 = sum r32 r32.
template <int NSP>
__global__ __launch_bounds__(256) void k_residual64(
    const int* __restrict__ rptr, const int* __restrict__ cptr, const double* __restrict__ are,
    const double* __restrict__ aim, const double* __restrict__ fscale,
    const double* __restrict__ b, const double* __restrict__ x64, float* __restrict__ r32,
    float* __restrict__ rhat32, float* __restrict__ p32, float* __restrict__ x32,
    const int* __restrict__ ostatus, int write, double* __restrict__ part1,
    double* __restrict__ part2, int nn, int ncb, int g0, int nbx, int tile_rows) {
  using G = Geo<double, NSP>;  // one complex128 per lane
  const int g = blockIdx.y, bx = blockIdx.x;
  __shared__ double sh[2 * 4 * NSP];
  __shared__ int srp[kChunkRows + 1];
  __shared__ int scol[kChunkCap];
  __shared__ double sval[2 * kChunkCap];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int c = lane % G::LPR;
  const long long rhs = (long long)g * NSP + c;
  const bool act = ostatus ? (ostatus[rhs] == ST_ACTIVE) : true;
  const int f = (g0 + g) / ncb;
  const size_t gbase = (size_t)g * nn;
  Acc<1> d1, d2;
  d1.zero();
  d2.zero();
  auto epi = [&](int row, const double* yr, const double* yi) {
    const size_t off = ((gbase + row) * NSP + c);
    const double2 bv = *reinterpret_cast<const double2*>(b + 2 * off);
    double rr = bv.x - yr[0], ri = bv.y - yi[0];
    if (!act) { rr = 0.0; ri = 0.0; }
    d1.re[0] += rr * rr + ri * ri;
    if (write) {
      const float fr = (float)rr, fi = (float)ri;
      d2.re[0] += (double)fr * fr - (double)fi * fi;
      d2.im[0] += 2.0 * (double)fr * fi;
      const float2 v = make_float2(fr, fi);
      reinterpret_cast<float2*>(r32)[off] = v;
      if (rhat32) reinterpret_cast<float2*>(rhat32)[off] = v;
      reinterpret_cast<float2*>(p32)[off] = v;
      reinterpret_cast<float2*>(x32)[off] = make_float2(0.f, 0.f);
    }
  };
  const int tile0 = xcd_tile(bx) * tile_rows;
  const int tile1 = min(tile0 + tile_rows, nn);
  if (x64) {
    spmv_rows<double, NSP, MAT_A, 8>(rptr, cptr, are, aim, fscale[f], x64 + (gbase * NSP + c) * 2, tile0, tile1,
                                     srp, scol, sval, epi);
  } else {
    const double z[1] = {0.0};
    for (int row = tile0 + wv * G::RPW + lane / G::LPR; row < tile1; row += G::RPB) epi(row, z, z);
  }
  block_reduce_store<double, NSP>(d1, sh, part1, (long long)g * nbx + bx);
  if (write) {
    __syncthreads();
    block_reduce_store<double, NSP>(d2, sh, part2, (long long)g * nbx + bx);
  }
}

// x64 += x32 for columns that took part in the sweep
template <int NSP>
__global__ __launch_bounds__(256) void k_add_correction(double* __restrict__ x64,
                                                        const float* __restrict__ x32,
                                                        const int* __restrict__ istatus, int nn) {
  const int g = blockIdx.y;
  const long long nper = (long long)nn * NSP;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nper;
       i += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(i % NSP);
    if (istatus[(long long)g * NSP + c] == ST_SKIP) continue;
    const long long o = (long long)g * nper + i;
    const float2 d = reinterpret_cast<const float2*>(x32)[o];
    double2 v = reinterpret_cast<double2*>(x64)[o];
    v.x += d.x;
    v.y += d.y;
    reinterpret_cast<double2*>(x64)[o] = v;
  }
}

// ------------------------------------------------ reference-structure passes
// part1 = sum a b (mode 0) or sum |a|^2 (mode 1)
template <typename T, int NSP>
__global__ __launch_bounds__(256) void k_dot(const T* __restrict__ a, const T* __restrict__ bb,
                                             int mode, double* __restrict__ part1, int nn, int nbx, int tile_rows,
                                             const int* __restrict__ grp_active) {
  Acc<Geo<T, NSP>::CPL> d1;
  d1.zero();
  ROW_LOOP_BEGIN
  T ar[CPL], ai[CPL], br[CPL], bi[CPL];
  Pack<T>::load(a + off, ar, ai);
  if (mode == 0) Pack<T>::load(bb + off, br, bi);
#pragma unroll
  for (int e = 0; e < CPL; ++e) {
    if (mode == 0) {
      d1.re[e] += (double)ar[e] * br[e] - (double)ai[e] * bi[e];
      d1.im[e] += (double)ar[e] * bi[e] + (double)ai[e] * br[e];
    } else {
      d1.re[e] += (double)ar[e] * ar[e] + (double)ai[e] * ai[e];
    }
  }
  ROW_LOOP_END
  block_reduce_store<T, NSP>(d1, sh, part1, (long long)g * nbx + bx);
}

// out = a + op(coef) b; op 0: 1/coef, 1: -1/coef, 2: coef, 3: -coef
template <typename T, int NSP>
__global__ __launch_bounds__(256) void k_axpby(T* __restrict__ out, const T* __restrict__ a,
                                               const T* __restrict__ bb,
                                               const double* __restrict__ coef, int op, int nn,
                                               int nbx, int tile_rows, const int* __restrict__ grp_active) {
  T cr[Geo<T, NSP>::CPL], ci[Geo<T, NSP>::CPL];
  ROW_LOOP_BEGIN
  if (st == 0) {
#pragma unroll
    for (int e = 0; e < CPL; ++e) {
      double xr = coef[2 * (rhs0 + e)], xi = coef[2 * (rhs0 + e) + 1];
      if (op <= 1) {
        const double d = xr * xr + xi * xi;
        xr = xr / d;
        xi = -xi / d;
      }
      if (op == 1 || op == 3) { xr = -xr; xi = -xi; }
      cr[e] = (T)xr;
      ci[e] = (T)xi;
    }
  }
  T ar[CPL], ai[CPL], br[CPL], bi[CPL];
  Pack<T>::load(a + off, ar, ai);
  Pack<T>::load(bb + off, br, bi);
#pragma unroll
  for (int e = 0; e < CPL; ++e) {
    ar[e] += cr[e] * br[e] - ci[e] * bi[e];
    ai[e] += cr[e] * bi[e] + ci[e] * br[e];
  }
  Pack<T>::store(out + off, ar, ai);
  ROW_LOOP_END
  (void)bx;
}

// ------------------------------------------------------------- scalars ----
struct BState {
  double* bnorm;   // nrhs
  double* relres;  // nrhs
  double* tol;     // nrhs
  double* rho;     // 2*nrhs
  double* alpha;   // 2*nrhs
  double* omega;   // 2*nrhs
  double* beta;    // 2*nrhs
  int* status;     // nrhs
  int* half;       // nrhs
  int* bd;         // nrhs
  int* iters;      // nrhs
  int* grp_active; // ngroups
  int* grp_prev;   // ngroups: active at the start of the current COCG iteration
  int* counter;    // 4
};

// Sum of the per-block partials of every column of group g.  All threads of
// the (kScalThreads-wide) block call it; thread c < nsp receives column c's
// sum.  Each column is reduced by kScalThreads/nsp threads in parallel.
constexpr int kScalThreads = 256;
__device__ inline void block_sum_part(const double* __restrict__ part, long long g, int nsp, int nbx,
                                      double& re, double& im) {
  __shared__ double sr[kScalThreads], si[kScalThreads];
  const int t = threadIdx.x, c = t % nsp, q = t / nsp, nq = kScalThreads / nsp;
  double a = 0.0, b = 0.0;
  for (int bx = q; bx < nbx; bx += nq) {
    const long long i = 2 * ((g * nbx + bx) * nsp + c);
    a += part[i];
    b += part[i + 1];
  }
  sr[t] = a;
  si[t] = b;
  __syncthreads();
  if (t < nsp) {
    for (int k = 1; k < nq; ++k) {
      a += sr[t + k * nsp];
      b += si[t + k * nsp];
    }
    re = a;
    im = b;
  }
  __syncthreads();
}

__device__ inline void group_flag(BState s, int g, bool active) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const unsigned long long m = __ballot(active);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(&cnt, __popcll(m));
  __syncthreads();
  if (threadIdx.x == 0) {
    const int n = cnt;
    s.grp_active[g] = n > 0;
    if (n) {
      atomicAdd(s.counter, n);
      atomicAdd(s.counter + 1, 1);
    }
  }
}

// init from (part1 = |b|^2, part2 = b.b); `mask` (optional outer status)
// leaves non-active columns skipped; columns >= ncols are padding.
__global__ __launch_bounds__(kScalThreads) void k_sc_init(BState s, const double* __restrict__ part1,
                                                          const double* __restrict__ part2,
                                                          const int* __restrict__ mask, int nsp,
                                                          int nbx, int ncb, int ncols, int g0) {
  const int g = blockIdx.x, c = threadIdx.x;
  double n2 = 0, dummy = 0, rr = 0, ri = 0;
  block_sum_part(part1, g, nsp, nbx, n2, dummy);
  block_sum_part(part2, g, nsp, nbx, rr, ri);
  bool active = false;
  if (c < nsp) {
    const long long i = (long long)g * nsp + c;
    const int col = ((g0 + g) % ncb) * nsp + c;
    s.iters[i] = 0;
    s.half[i] = 0;
    s.bd[i] = 0;
    s.bnorm[i] = sqrt(n2);
    s.rho[2 * i] = rr;
    s.rho[2 * i + 1] = ri;
    s.alpha[2 * i] = s.alpha[2 * i + 1] = 0.0;
    s.omega[2 * i] = s.omega[2 * i + 1] = 0.0;
    s.beta[2 * i] = s.beta[2 * i + 1] = 0.0;
    if (mask && mask[i] != ST_ACTIVE) {
      s.status[i] = ST_SKIP;
    } else if (col >= ncols || n2 == 0.0) {
      s.status[i] = DOGPUP_RHS_ZERO;
      s.relres[i] = 0.0;
    } else {
      s.status[i] = ST_ACTIVE;
      s.relres[i] = 1.0;
      active = true;
    }
  }
  group_flag(s, g, active);
}

__global__ __launch_bounds__(kScalThreads) void k_sc_alpha(BState s, const double* __restrict__ part1,
                                                           int nsp, int nbx) {
  const int g = blockIdx.x, c = threadIdx.x;
  if (!s.grp_active[g]) return;
  double dr = 0, di = 0;
  block_sum_part(part1, g, nsp, nbx, dr, di);
  if (c >= nsp) return;
  const long long i = (long long)g * nsp + c;
  if (s.status[i] != ST_ACTIVE) return;
  if ((dr == 0.0 && di == 0.0) || !isfinite(dr) || !isfinite(di)) {
    s.status[i] = DOGPUP_RHS_BREAKDOWN;
    return;
  }
  const cpx<double> a = cdiv(cpx<double>{s.rho[2 * i], s.rho[2 * i + 1]}, cpx<double>{dr, di});
  s.alpha[2 * i] = a.re;
  s.alpha[2 * i + 1] = a.im;
}

__global__ __launch_bounds__(kScalThreads) void k_sc_s(BState s, const double* __restrict__ part1, int nsp,
                                                       int nbx) {
  const int g = blockIdx.x, c = threadIdx.x;
  if (!s.grp_active[g]) return;
  double n2 = 0, dummy = 0;
  block_sum_part(part1, g, nsp, nbx, n2, dummy);
  if (c >= nsp) return;
  const long long i = (long long)g * nsp + c;
  if (s.status[i] != ST_ACTIVE) return;
  const double rs = sqrt(n2) / s.bnorm[i];
  if (rs < s.tol[i]) {
    s.half[i] = 1;
    s.relres[i] = rs;
  }
}

__global__ __launch_bounds__(kScalThreads) void k_sc_omega(BState s, const double* __restrict__ part1,
                                                           const double* __restrict__ part2, int nsp,
                                                           int nbx) {
  const int g = blockIdx.x, c = threadIdx.x;
  if (!s.grp_active[g]) return;
  double nr = 0, ni = 0, dr = 0, di = 0;
  block_sum_part(part1, g, nsp, nbx, nr, ni);
  block_sum_part(part2, g, nsp, nbx, dr, di);
  if (c >= nsp) return;
  const long long i = (long long)g * nsp + c;
  if (s.status[i] != ST_ACTIVE) return;
  double om_r = 0.0, om_i = 0.0;
  if (!s.half[i]) {
    if ((dr == 0.0 && di == 0.0) || !isfinite(dr) || !isfinite(di) || !isfinite(nr) || !isfinite(ni)) {
      s.bd[i] = 1;
    } else {
      const cpx<double> o = cdiv(cpx<double>{nr, ni}, cpx<double>{dr, di});
      om_r = o.re;
      om_i = o.im;
    }
  }
  s.omega[2 * i] = om_r;
  s.omega[2 * i + 1] = om_i;
}

__global__ __launch_bounds__(kScalThreads) void k_sc_beta(BState s, const double* __restrict__ part1,
                                                          const double* __restrict__ part2, int nsp,
                                                          int nbx, int maxit) {
  const int g = blockIdx.x, c = threadIdx.x;
  if (!s.grp_active[g]) return;   // inactive groups stay inactive
  double n2 = 0, dummy = 0, rr = 0, ri = 0;
  block_sum_part(part2, g, nsp, nbx, n2, dummy);
  block_sum_part(part1, g, nsp, nbx, rr, ri);
  bool active = false;
  if (c < nsp) {
    const long long i = (long long)g * nsp + c;
    if (s.status[i] == ST_ACTIVE) {
      const int it = ++s.iters[i];
      const double rn = sqrt(n2) / s.bnorm[i];
      if (s.half[i]) {
        s.status[i] = DOGPUP_RHS_CONVERGED;
      } else {
        s.relres[i] = rn;
        const double om_r = s.omega[2 * i], om_i = s.omega[2 * i + 1];
        if (s.bd[i] || !isfinite(rn)) {
          s.status[i] = DOGPUP_RHS_BREAKDOWN;
        } else if (rn < s.tol[i]) {
          s.status[i] = DOGPUP_RHS_CONVERGED;
        } else if (it >= maxit) {
          s.status[i] = DOGPUP_RHS_MAXIT;
        } else if ((rr == 0.0 && ri == 0.0) || (om_r == 0.0 && om_i == 0.0)) {
          s.status[i] = DOGPUP_RHS_BREAKDOWN;
        } else {
          const cpx<double> q = cdiv(cpx<double>{rr, ri}, cpx<double>{s.rho[2 * i], s.rho[2 * i + 1]});
          const cpx<double> a_over_w =
              cdiv(cpx<double>{s.alpha[2 * i], s.alpha[2 * i + 1]}, cpx<double>{om_r, om_i});
          const cpx<double> b = cmul(q, a_over_w);
          s.beta[2 * i] = b.re;
          s.beta[2 * i + 1] = b.im;
          s.rho[2 * i] = rr;
          s.rho[2 * i + 1] = ri;
          active = true;
        }
      }
    }
  }
  group_flag(s, g, active);
}

// ---- COCG scalars (conjugate orthogonal CG: CG with the unconjugated
// bilinear form, for complex-symmetric A and the symmetric M^-1 = G^T G).
// rho = (G r).(G r) = r.M^-1 r, alpha = rho / (p.A p), beta = rho_new / rho.
// A column that is not active gets alpha = beta = 0, so its vectors stay
// bounded while the rest of its group iterates.

// After the start pass u = G r (part1 = u.u): rho = u.u, beta = 0.
__global__ __launch_bounds__(kScalThreads) void k_cg_start(BState s, const double* __restrict__ part1,
                                                           int nsp, int nbx) {
  const int g = blockIdx.x, c = threadIdx.x;
  double ur = 0, ui = 0;
  block_sum_part(part1, g, nsp, nbx, ur, ui);
  bool active = false;
  if (c < nsp) {
    const long long i = (long long)g * nsp + c;
    if (s.status[i] == ST_ACTIVE) {
      if ((ur == 0.0 && ui == 0.0) || !isfinite(ur) || !isfinite(ui)) {
        s.status[i] = DOGPUP_RHS_BREAKDOWN;
      } else {
        s.rho[2 * i] = ur;
        s.rho[2 * i + 1] = ui;
        active = true;
      }
    }
    s.alpha[2 * i] = s.alpha[2 * i + 1] = 0.0;
    s.beta[2 * i] = s.beta[2 * i + 1] = 0.0;
  }
  group_flag(s, g, active);
}

// part1 = p.(A p)
__global__ __launch_bounds__(kScalThreads) void k_cg_alpha(BState s, const double* __restrict__ part1,
                                                           int nsp, int nbx) {
  const int g = blockIdx.x, c = threadIdx.x;
  if (!s.grp_active[g]) return;
  double dr = 0, di = 0;
  block_sum_part(part1, g, nsp, nbx, dr, di);
  if (c >= nsp) return;
  const long long i = (long long)g * nsp + c;
  double ar = 0.0, ai = 0.0;
  if (s.status[i] == ST_ACTIVE) {
    if ((dr == 0.0 && di == 0.0) || !isfinite(dr) || !isfinite(di)) {
      s.status[i] = DOGPUP_RHS_BREAKDOWN;
    } else {
      const cpx<double> a = cdiv(cpx<double>{s.rho[2 * i], s.rho[2 * i + 1]}, cpx<double>{dr, di});
      ar = a.re;
      ai = a.im;
    }
  }
  s.alpha[2 * i] = ar;
  s.alpha[2 * i + 1] = ai;
}

// part2 = |r|^2 (after the update), part1 = u.u with u = G r: convergence,
// then beta and rho for the columns that go on.
__global__ __launch_bounds__(kScalThreads) void k_cg_beta(BState s, const double* __restrict__ part1,
                                                          const double* __restrict__ part2, int nsp,
                                                          int nbx, int maxit) {
  const int g = blockIdx.x, c = threadIdx.x;
  // grp_prev gates this iteration's G^T pass, which applies x += alpha p
  const int was = s.grp_active[g];
  if (c == 0) s.grp_prev[g] = was;
  if (!was) return;
  double n2 = 0, dummy = 0, ur = 0, ui = 0;
  block_sum_part(part2, g, nsp, nbx, n2, dummy);
  block_sum_part(part1, g, nsp, nbx, ur, ui);
  bool active = false;
  if (c < nsp) {
    const long long i = (long long)g * nsp + c;
    double br = 0.0, bi = 0.0;
    if (s.status[i] == ST_ACTIVE) {
      const int it = ++s.iters[i];
      const double rn = sqrt(n2) / s.bnorm[i];
      s.relres[i] = rn;
      if (!isfinite(rn)) s.status[i] = DOGPUP_RHS_BREAKDOWN;
      else if (rn < s.tol[i]) s.status[i] = DOGPUP_RHS_CONVERGED;
      else if (it >= maxit) s.status[i] = DOGPUP_RHS_MAXIT;
      else if ((ur == 0.0 && ui == 0.0) || !isfinite(ur) || !isfinite(ui)) s.status[i] = DOGPUP_RHS_BREAKDOWN;
      else {
        const cpx<double> b = cdiv(cpx<double>{ur, ui}, cpx<double>{s.rho[2 * i], s.rho[2 * i + 1]});
        br = b.re;
        bi = b.im;
        s.rho[2 * i] = ur;
        s.rho[2 * i + 1] = ui;
        active = true;
      }
    }
    s.beta[2 * i] = br;
    s.beta[2 * i + 1] = bi;
  }
  group_flag(s, g, active);
}

// Outer (FP64) refinement state
struct RState {
  double* bnorm;    // nrhs
  double* relres;   // nrhs
  double* prev;     // nrhs
  int* status;      // nrhs
  int* iters;       // nrhs
};

// After a residual pass: decide each column's outer status and its inner
// tolerance.  part1 = |r64|^2.
__global__ __launch_bounds__(kScalThreads) void k_rf_check(RState o, BState in,
                                                           const double* __restrict__ part1, int nsp,
                                                           int nbx, int ncb, int ncols, int g0,
                                                           int sweep, double tol, double tol_in,
                                                           int maxit) {
  const int g = blockIdx.x, c = threadIdx.x;
  double n2 = 0, dummy = 0;
  block_sum_part(part1, g, nsp, nbx, n2, dummy);
  bool active = false;
  if (c < nsp) {
    const long long i = (long long)g * nsp + c;
    const int col = ((g0 + g) % ncb) * nsp + c;
    if (sweep == 0) {
      o.bnorm[i] = sqrt(n2);
      o.iters[i] = 0;
      o.prev[i] = INFINITY;
      o.status[i] = (col >= ncols || n2 == 0.0) ? DOGPUP_RHS_ZERO : ST_ACTIVE;
      o.relres[i] = (o.status[i] == ST_ACTIVE) ? 1.0 : 0.0;
    }
    if (o.status[i] == ST_ACTIVE) {
      const double rel = sqrt(n2) / o.bnorm[i];
      o.relres[i] = rel;
      if (rel < tol) o.status[i] = DOGPUP_RHS_CONVERGED;
      else if (!isfinite(rel)) o.status[i] = DOGPUP_RHS_BREAKDOWN;
      else if (o.iters[i] >= maxit) o.status[i] = DOGPUP_RHS_MAXIT;
      else if (rel > 0.5 * o.prev[i]) o.status[i] = DOGPUP_RHS_BREAKDOWN;  // stagnation
      else {
        o.prev[i] = rel;
        double t = 0.5 * tol / rel;
        if (t < tol_in) t = tol_in;
        if (t > 0.5) t = 0.5;
        in.tol[i] = t;
        active = true;
      }
    }
  }
  group_flag(in, g, active);
}

__global__ void k_rf_accum(RState o, BState in, int nsp) {
  const int g = blockIdx.x, c = threadIdx.x;
  if (c >= nsp) return;
  const long long i = (long long)g * nsp + c;
  if (in.status[i] != ST_SKIP) o.iters[i] += in.iters[i];
}

// Reference-structure scalars (lockstep).  op 0: alph = (v.rhat0)/rho0;
// op 1: err = |s|/err0 -> max; op 2: om = (Gt.Gs)/(Gt.Gt); op 3: err =
// |r|/err0 -> max; op 4: rho = rhat0.r, beta = rho/rho0 / (alph om), rho0 =
// rho; op 5: init err0 = |r|, rho0 = rhat0.r.
__global__ __launch_bounds__(kScalThreads) void k_plain_scal(BState s, int op,
                                                             const double* __restrict__ part1,
                                                             const double* __restrict__ part2, int nsp,
                                                             int nbx, int ncb, int ncols, int g0,
                                                             unsigned long long* __restrict__ maxbits) {
  const int g = blockIdx.x, c = threadIdx.x;
  double ar = 0, ai = 0, br = 0, bi = 0;
  block_sum_part(part1, g, nsp, nbx, ar, ai);
  if (part2) block_sum_part(part2, g, nsp, nbx, br, bi);
  if (c >= nsp) return;
  const long long i = (long long)g * nsp + c;
  const int col = ((g0 + g) % ncb) * nsp + c;
  const bool valid = col < ncols && s.bnorm[i] > 0.0;
  switch (op) {
    case 5:
      s.bnorm[i] = sqrt(ar);
      s.rho[2 * i] = br;
      s.rho[2 * i + 1] = bi;
      s.iters[i] = 0;
      s.status[i] = (col >= ncols || ar == 0.0) ? DOGPUP_RHS_ZERO : ST_ACTIVE;
      s.relres[i] = s.status[i] == ST_ACTIVE ? 1.0 : 0.0;
      break;
    case 0: {
      cpx<double> a{0.0, 0.0};
      if (valid) a = cdiv(cpx<double>{ar, ai}, cpx<double>{s.rho[2 * i], s.rho[2 * i + 1]});
      if (!valid || (a.re == 0.0 && a.im == 0.0)) a = cpx<double>{INFINITY, 0.0};
      s.alpha[2 * i] = a.re;
      s.alpha[2 * i + 1] = a.im;
      break;
    }
    case 1:
    case 3: {
      double e = valid ? sqrt(ar) / s.bnorm[i] : 0.0;
      if (!(e == e)) e = INFINITY;
      s.relres[i] = e;
      if (valid && op == 3) s.iters[i] += 1;
      atomicMax(maxbits, (unsigned long long)__double_as_longlong(e));
      break;
    }
    case 2: {
      cpx<double> o{0.0, 0.0};
      if (valid && (br != 0.0 || bi != 0.0)) o = cdiv(cpx<double>{ar, ai}, cpx<double>{br, bi});
      s.omega[2 * i] = o.re;
      s.omega[2 * i + 1] = o.im;
      break;
    }
    case 4: {
      cpx<double> b{0.0, 0.0};
      if (valid) {
        const cpx<double> rho0{s.rho[2 * i], s.rho[2 * i + 1]};
        const cpx<double> ao =
            cmul(cpx<double>{s.alpha[2 * i], s.alpha[2 * i + 1]}, cpx<double>{s.omega[2 * i], s.omega[2 * i + 1]});
        b = cdiv(cdiv(cpx<double>{ar, ai}, rho0), ao);
        if (!isfinite(b.re) || !isfinite(b.im)) b = cpx<double>{0.0, 0.0};
        s.rho[2 * i] = ar;
        s.rho[2 * i + 1] = ai;
      }
      s.beta[2 * i] = b.re;
      s.beta[2 * i + 1] = b.im;
      break;
    }
  }
}

// ------------------------------------------------------------ the handle --
struct Buf {
  void* p = nullptr;
  size_t n = 0;
  ~Buf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  int ensure(size_t bytes) {
    if (bytes <= n && p) return DOGPUP_OK;
    release();
    hipError_t e = hipMalloc(&p, bytes ? bytes : 16);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(DOGPUP_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    n = bytes;
    return DOGPUP_OK;
  }
  template <typename T>
  T* as(size_t off_bytes = 0) const { return reinterpret_cast<T*>(reinterpret_cast<char*>(p) + off_bytes); }
};

#define DP_TRY(x)                         \
  do {                                    \
    int rc_ = (x);                        \
    if (rc_ != DOGPUP_OK) return rc_;     \
  } while (0)

}  // namespace dogpup

struct dogpup_solver {
  int nn = 0, nf = 0, ncols = 0, nsp = 0, ncb = 0, ngroups = 0, mode = 0;
  int nnzA = 0, nnzG = 0;
  int method = DOGPUP_METHOD_BICGSTAB;  // Krylov method of the fused forms
  // caller's patterns in the reference node order (FSAI setup runs on these)
  const int *rA0 = nullptr, *cA0 = nullptr, *rG0 = nullptr, *cG0 = nullptr;
  // patterns in solver node order (owned): A, G, G^T and the value maps
  // slot -> slot of the reference-order arrays
  const int *rA = nullptr, *cA = nullptr, *rG = nullptr, *cG = nullptr;
  const int *rGT = nullptr, *cGT = nullptr;
  dogpup::Buf pAr, pAc, pGr, pGc, pGTr, pGTc, amap, gmap, gtmap, order;
  int unroll_a = 8, unroll_g = 8;  // gathers in flight per row (k_spmv: 4, 8 or 16)
  // values: a64re/a64im and g64 in reference order; the rest in solver order
  dogpup::Buf a64re, a64im, a64pre, a64pim, a32re, a32im, fscale, g64, gp64, gtp64, g32, gt32;
  dogpup::Buf ws, part1, part2, state, rstate, bperm, xperm;
  int chunk = 0, nbx = 0, tile_rows = 512, lds_pad = 0;
  bool have_matrix = false;
  double stats[4] = {0, 0, 0, 0};
  int* h_counter = nullptr;  // pinned: [0] active RHS, [1] active groups
  unsigned long long* h_max = nullptr;
  int active_groups = 0;
  // optional per-launch timing: event pairs resolved after the solve
  struct Rec {
    hipEvent_t a, b;
    int cls;
    double bytes;
  };
  bool profile = false;
  std::vector<Rec> recs;
  // captured BiCGSTAB iterations, keyed by (chunk, maxit, precision)
  bool use_graph = true;
  int graph_k = 4;
  std::map<long long, hipGraphExec_t> graphs;
  hipStream_t own = nullptr;   // the solve's own (capturable) stream
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  void clear_graphs() {
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
  }
  double prof_ms[4] = {0, 0, 0, 0}, prof_bytes[4] = {0, 0, 0, 0}, prof_n[4] = {0, 0, 0, 0};
  ~dogpup_solver() {
    clear_graphs();
    if (own) (void)hipStreamDestroy(own);
    if (ev_in) (void)hipEventDestroy(ev_in);
    if (ev_out) (void)hipEventDestroy(ev_out);
    for (auto& r : recs) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    if (h_counter) (void)hipHostFree(h_counter);
    if (h_max) (void)hipHostFree(h_max);
  }
};

namespace dogpup {

__global__ void k_cast_f32(const double* __restrict__ a, float* __restrict__ b, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// out[p] = in[map[p]] (real values, type conversion)
template <typename TO>
__global__ void k_gather_real(const double* __restrict__ in, const int* __restrict__ map,
                              TO* __restrict__ out, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = (TO)in[map[i]];
}

// out[f][p] = in[f][map[p]] (complex [f][nnz])
template <typename TO>
__global__ void k_gather_cplx(const double* __restrict__ in, const int* __restrict__ map,
                              TO* __restrict__ out, long long nnz, int nf) {
  const long long total = nnz * nf;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const long long f = i / nnz, p = i % nnz;
    const long long s = f * nnz + map[p];
    out[2 * i] = (TO)in[2 * s];
    out[2 * i + 1] = (TO)in[2 * s + 1];
  }
}

// Row permutation of a complex128 block [g][node][nsp]: dir 0 gathers into
// solver order (dst[g][i] = src[g][order[i]]), dir 1 scatters back.
__global__ void k_permute_rows(const double* __restrict__ src, double* __restrict__ dst,
                               const int* __restrict__ order, int nn, int nsp, int gc, int dir) {
  const long long per = (long long)nn * nsp;
  const long long total = per * gc;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const long long g = i / per, r = i % per;
    const int node = (int)(r / nsp), c = (int)(r % nsp);
    const long long o = g * per + (long long)order[node] * nsp + c;
    const double2 v = reinterpret_cast<const double2*>(src)[dir == 0 ? o : i];
    reinterpret_cast<double2*>(dst)[dir == 0 ? i : o] = v;
  }
}


static unsigned grid1(long long n) {
  long long b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// Clears n ints (a kernel, so that captured graphs hold no memset node).
__global__ void k_zero_ints(int* __restrict__ p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// Everything one chunk of groups needs on the device.
template <typename T>
struct Vecs {
  T *x, *r, *rhat, *p, *v, *h, *s, *z, *t, *u, *w;
};

template <typename T, int NSP>
struct Chunk {
  dogpup_solver* S;
  hipStream_t st;
  int gcount, g0;
  BState bs;
  double *p1, *p2;

  dim3 vgrid() const { return dim3(S->nbx, gcount); }
  const T* a0() const { return sizeof(T) == 4 ? (const T*)S->a32re.p : (const T*)S->a64pre.p; }
  const T* a1() const { return sizeof(T) == 4 ? (const T*)S->a32im.p : (const T*)S->a64pim.p; }
  const T* gvals(bool tr) const {
    if (sizeof(T) == 4) return tr ? (const T*)S->gt32.p : (const T*)S->g32.p;
    return tr ? (const T*)S->gtp64.p : (const T*)S->gp64.p;
  }

  template <int MAT, int EPI, int U>
  int pad_lds() {
    if (S->lds_pad > 0)
      DP_HIP(hipFuncSetAttribute((const void*)k_spmv<T, NSP, MAT, EPI, U>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, S->lds_pad));
    return DOGPUP_OK;
  }
  template <int MAT, int EPI, int U>
  int launch(const int* rp, const int* cp, const T* m0, const T* m1, long long nz, const T* x, T* y,
             const T* w, const int* ga, const double* coef, const double* coef2, T* xacc) {
    DP_TRY((pad_lds<MAT, EPI, U>()));
    k_spmv<T, NSP, MAT, EPI, U><<<vgrid(), 256, S->lds_pad, st>>>(
        rp, cp, m0, m1, nz, S->fscale.template as<double>(), x, y, w, coef, coef2, xacc, p1, p2, S->nn, S->ncb,
        g0, S->nbx, S->tile_rows, ga);
    DP_LAUNCH_CHECK();
    return DOGPUP_OK;
  }
  template <int MAT, int EPI>
  int launch_u(int u, const int* rp, const int* cp, const T* m0, const T* m1, long long nz, const T* x, T* y,
               const T* w, const int* ga, const double* coef, const double* coef2, T* xacc) {
    switch (u) {
      case 4: return launch<MAT, EPI, 4>(rp, cp, m0, m1, nz, x, y, w, ga, coef, coef2, xacc);
      case 8: return launch<MAT, EPI, 8>(rp, cp, m0, m1, nz, x, y, w, ga, coef, coef2, xacc);
      default: return launch<MAT, EPI, 16>(rp, cp, m0, m1, nz, x, y, w, ga, coef, coef2, xacc);
    }
  }
  template <int EPI>
  int spmvA(const T* x, T* y, const T* w, const int* ga) {
    pbegin();
    DP_TRY((launch_u<MAT_A, EPI>(S->unroll_a, S->rA, S->cA, a0(), a1(), 0, x, y, w, ga, nullptr, nullptr,
                                 nullptr)));
    pend(1, groups(ga != nullptr) * vbytes() * ((EPI == EPI_STORE || w == x) ? 2 : 3) +
                (double)S->nnzA * (2 * sizeof(T) + 4) + 4.0 * (S->nn + 1));
    return DOGPUP_OK;
  }
  template <int EPI>
  int spmvG(bool tr, const T* x, T* y, const T* w, const int* ga, const double* coef = nullptr,
            const double* coef2 = nullptr, T* xacc = nullptr) {
    pbegin();
    DP_TRY((launch_u<MAT_G, EPI>(S->unroll_g, tr ? S->rGT : S->rG, tr ? S->cGT : S->cG, gvals(tr), nullptr,
                                 S->nnzG, x, y, w, ga, coef, coef2, xacc)));
    const int streams = (EPI == EPI_STORE || EPI == EPI_DOT2 || EPI == EPI_STORE_SELF) ? 2
                        : (EPI == EPI_AXPBY_X ? 5 : 3);
    pend(0, groups(ga != nullptr) * (vbytes() * streams + (double)S->nnzG * (2 * sizeof(T) + 4) +
                                    4.0 * (S->nn + 1)));
    return DOGPUP_OK;
  }

  int wait_count(int* out) {
    DP_HIP(hipMemcpyAsync(S->h_counter, bs.counter, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    DP_HIP(hipStreamSynchronize(st));
    *out = S->h_counter[0];
    S->active_groups = S->h_counter[1];
    return DOGPUP_OK;
  }
  int reset_count() {
    k_zero_ints<<<1, 64, 0, st>>>(bs.counter, 2);
    DP_LAUNCH_CHECK();
    return DOGPUP_OK;
  }
  // profiling brackets: class 0 = G/G^T SpMV, 1 = A SpMV, 2 = vector passes,
  // 3 = FP64 residual / correction
  hipEvent_t pa = nullptr;
  void pbegin() {
    if (!S->profile) return;
    (void)hipEventCreate(&pa);
    (void)hipEventRecord(pa, st);
  }
  void pend(int cls, double bytes) {
    if (!S->profile) return;
    hipEvent_t pb;
    (void)hipEventCreate(&pb);
    (void)hipEventRecord(pb, st);
    S->recs.push_back({pa, pb, cls, bytes});
  }
  double vbytes() const { return (double)S->nn * NSP * 2 * sizeof(T); }
  double groups(bool skip) const { return skip ? (double)S->active_groups : (double)gcount; }

  // Fused BiCGSTAB.  If b != nullptr the init pass runs from b (x = 0);
  // otherwise r, rhat, p, x and the state are already initialised (the
  // refinement residual pass did it).  Returns iterations run.
  // One BiCGSTAB iteration over every active group (no host synchronisation,
  // no memset nodes: it is captured into a HIP graph when profiling is off).
  int body(Vecs<T> V, int maxit) {
    const int* ga = bs.grp_active;
    const int nsp = NSP;
    DP_TRY(spmvG<EPI_STORE>(false, V.p, V.u, nullptr, ga));      // u = G p
    DP_TRY(spmvG<EPI_STORE>(true, V.u, V.h, nullptr, ga));       // h = G^T u
    DP_TRY(spmvA<EPI_STORE_DOT>(V.h, V.v, V.rhat, ga));          // v = A h, rhat.v
    k_sc_alpha<<<gcount, kScalThreads, 0, st>>>(bs, p1, nsp, S->nbx);
    DP_LAUNCH_CHECK();
    pbegin();
    k_axpy_s<T, NSP><<<vgrid(), 256, 0, st>>>(V.r, V.v, V.s, bs.alpha, p1, S->nn, S->nbx, S->tile_rows, ga);
    DP_LAUNCH_CHECK();
    pend(2, groups(true) * vbytes() * 3);
    k_sc_s<<<gcount, kScalThreads, 0, st>>>(bs, p1, nsp, S->nbx);
    DP_LAUNCH_CHECK();
    DP_TRY(spmvG<EPI_STORE>(false, V.s, V.u, nullptr, ga));      // u = G s
    DP_TRY(spmvG<EPI_STORE>(true, V.u, V.z, nullptr, ga));       // z = G^T G s
    DP_TRY(spmvA<EPI_STORE>(V.z, V.t, nullptr, ga));             // t = A z
    DP_TRY(spmvG<EPI_DOT2>(false, V.t, nullptr, V.u, ga));       // (Gt.Gs), (Gt.Gt)
    k_sc_omega<<<gcount, kScalThreads, 0, st>>>(bs, p1, p2, nsp, S->nbx);
    DP_LAUNCH_CHECK();
    pbegin();
    k_update_xr<T, NSP><<<vgrid(), 256, 0, st>>>(V.x, V.h, V.z, V.r, V.s, V.t, V.rhat, bs.alpha,
                                                bs.omega, bs.status, p1, p2, S->nn, S->nbx, S->tile_rows, ga);
    DP_LAUNCH_CHECK();
    pend(2, groups(true) * vbytes() * 8);
    DP_TRY(reset_count());
    k_sc_beta<<<gcount, kScalThreads, 0, st>>>(bs, p1, p2, nsp, S->nbx, maxit);
    DP_LAUNCH_CHECK();
    const int ag = S->active_groups;  // groups still active before this update
    pbegin();
    k_update_p<T, NSP><<<vgrid(), 256, 0, st>>>(V.p, V.r, V.v, bs.beta, bs.omega, S->nn, S->nbx, S->tile_rows, ga);
    DP_LAUNCH_CHECK();
    pend(2, (double)ag * vbytes() * 4);
    return DOGPUP_OK;
  }

  // One COCG iteration over every active group: q = A p (p.q), alpha,
  // r -= alpha q (|r|^2), u = G r (u.u), convergence and beta, then
  // p = G^T u + beta p with x += alpha p (old p) in the same pass.  Three
  // operator applications and 12 vector streams per iteration (BiCGSTAB:
  // seven and 30).  The G^T pass runs for the groups that were active when
  // the iteration began (grp_prev), so a group that converges in this
  // iteration still receives its last x update.
  int body_cocg(Vecs<T> V, int maxit) {
    const int* ga = bs.grp_active;
    const int nsp = NSP;
    DP_TRY(spmvA<EPI_STORE_DOT>(V.p, V.v, V.p, ga));             // q = A p, p.q
    k_cg_alpha<<<gcount, kScalThreads, 0, st>>>(bs, p1, nsp, S->nbx);
    DP_LAUNCH_CHECK();
    pbegin();
    k_cg_update<T, NSP><<<vgrid(), 256, 0, st>>>(V.r, V.v, bs.alpha, p2, S->nn, S->nbx, S->tile_rows, ga);
    DP_LAUNCH_CHECK();
    pend(2, groups(true) * vbytes() * 3);
    DP_TRY(spmvG<EPI_STORE_SELF>(false, V.r, V.u, nullptr, ga));  // u = G r, u.u
    DP_TRY(reset_count());
    k_cg_beta<<<gcount, kScalThreads, 0, st>>>(bs, p1, p2, nsp, S->nbx, maxit);
    DP_LAUNCH_CHECK();
    // p = G^T u + beta p, x += alpha p
    DP_TRY(spmvG<EPI_AXPBY_X>(true, V.u, V.p, nullptr, bs.grp_prev, bs.beta, bs.alpha, V.x));
    return DOGPUP_OK;
  }

  // COCG start from r (state initialised, x = 0): u = G r, rho = u.u,
  // p = G^T u (p holds finite values, beta = 0).
  int cocg_start(Vecs<T> V) {
    const int* ga = bs.grp_active;
    DP_TRY(spmvG<EPI_STORE_SELF>(false, V.r, V.u, nullptr, ga));
    DP_TRY(reset_count());
    k_cg_start<<<gcount, kScalThreads, 0, st>>>(bs, p1, NSP, S->nbx);
    DP_LAUNCH_CHECK();
    DP_TRY(spmvG<EPI_AXPBY>(true, V.u, V.p, nullptr, ga, bs.beta));
    return DOGPUP_OK;
  }

  int step(Vecs<T> V, int maxit) { return S->method == DOGPUP_METHOD_COCG ? body_cocg(V, maxit) : body(V, maxit); }

  // Fused BiCGSTAB or COCG.  If b != nullptr the init pass runs from b
  // (x = 0); otherwise r, rhat, p, x and the state are already initialised
  // (the refinement residual pass did it).  iters_run = loop iterations
  // issued.
  int run_fused(Vecs<T> V, const T* b, int maxit, const int* mask, int* iters_run) {
    const int nsp = NSP;
    if (b) {
      k_init<T, NSP><<<vgrid(), 256, 0, st>>>(b, V.x, V.r, V.rhat, V.p, p1, p2, S->nn, S->nbx, S->tile_rows, nullptr);
      DP_LAUNCH_CHECK();
      DP_TRY(reset_count());
      k_sc_init<<<gcount, kScalThreads, 0, st>>>(bs, p1, p2, mask, nsp, S->nbx, S->ncb, S->ncols, g0);
      DP_LAUNCH_CHECK();
    }
    if (S->method == DOGPUP_METHOD_COCG) DP_TRY(cocg_start(V));
    int active = 0;
    DP_TRY(wait_count(&active));
    int it = 0;
    if (S->profile || !S->use_graph) {
      while (active > 0 && it < maxit) {
        DP_TRY(step(V, maxit));
        ++it;
        DP_TRY(wait_count(&active));
      }
    } else {
      // Replay a captured iteration K times between convergence polls;
      // converged groups make every kernel exit early, so replays past
      // convergence cost only launch latency.
      const long long key = ((long long)g0 << 32) ^ ((long long)maxit << 3) ^ ((long long)S->method << 2) ^
                            (sizeof(T) == 4 ? 1 : 2);
      hipGraphExec_t ex = nullptr;
      auto found = S->graphs.find(key);
      if (found != S->graphs.end()) {
        ex = found->second;
      } else {
        DP_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = step(V, maxit);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(st, &g);
        if (rc != DOGPUP_OK) {
          if (g) (void)hipGraphDestroy(g);
          return rc;
        }
        if (ec != hipSuccess) return fail(DOGPUP_ERR_HIP, "graph capture: %s", hipGetErrorString(ec));
        const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) return fail(DOGPUP_ERR_HIP, "graph instantiate: %s", hipGetErrorString(ei));
        S->graphs[key] = ex;
      }
      const int K = S->graph_k;
      while (active > 0 && it < maxit) {
        for (int q = 0; q < K; ++q) DP_HIP(hipGraphLaunch(ex, st));
        it += K;
        DP_TRY(wait_count(&active));
      }
    }
    *iters_run = it;
    return DOGPUP_OK;
  }
};

}  // namespace dogpup

using namespace dogpup;

// Workspace carving
namespace dogpup {

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Vector slots carved from the workspace: x (only when the solver owns it,
// i.e. the complex64 correction), r, rhat, p, v, h, s, z, t, u (u also holds
// Gs), and w (only the reference-structure form materialises G t).
template <typename T>
static Vecs<T> carve(const Buf& ws, size_t vbytes, bool own_x, bool with_w) {
  Vecs<T> v{};
  size_t off = 0;
  auto take = [&]() { T* p = ws.as<T>(off); off += align256(vbytes); return p; };
  v.x = own_x ? take() : nullptr;
  v.r = take(); v.rhat = take(); v.p = take(); v.v = take(); v.h = take();
  v.s = take(); v.z = take(); v.t = take(); v.u = take();
  v.w = with_w ? take() : nullptr;
  return v;
}
static int nvec_for(int mode) { return mode == DOGPUP_MODE_C64_REFINE ? 10 : (mode == DOGPUP_MODE_C128_PLAIN ? 10 : 9); }

static BState carve_state(const Buf& b, int nrhs, int ngroups) {
  BState s;
  size_t off = 0;
  auto d = [&](int n) { double* p = b.as<double>(off); off += align256((size_t)n * 8); return p; };
  auto i = [&](int n) { int* p = b.as<int>(off); off += align256((size_t)n * 4); return p; };
  s.bnorm = d(nrhs); s.relres = d(nrhs); s.tol = d(nrhs);
  s.rho = d(2 * nrhs); s.alpha = d(2 * nrhs); s.omega = d(2 * nrhs); s.beta = d(2 * nrhs);
  s.status = i(nrhs); s.half = i(nrhs); s.bd = i(nrhs); s.iters = i(nrhs);
  s.grp_active = i(ngroups); s.grp_prev = i(ngroups); s.counter = i(4);
  return s;
}
static size_t state_bytes(int nrhs, int ngroups) {
  return 7 * align256((size_t)nrhs * 16) + 4 * align256((size_t)nrhs * 4) + 2 * align256((size_t)ngroups * 4) + 256;
}
static RState carve_rstate(const Buf& b, int nrhs) {
  RState s;
  size_t off = 0;
  auto d = [&](int n) { double* p = b.as<double>(off); off += align256((size_t)n * 8); return p; };
  auto i = [&](int n) { int* p = b.as<int>(off); off += align256((size_t)n * 4); return p; };
  s.bnorm = d(nrhs); s.relres = d(nrhs); s.prev = d(nrhs);
  s.status = i(nrhs); s.iters = i(nrhs);
  return s;
}

__global__ void k_fill(double* __restrict__ a, double v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// Copy per-RHS results of one chunk into host arrays indexed freq*ncols+col.
static int export_results(dogpup_solver* S, hipStream_t st, int g0, int gcount, const int* d_iters,
                          const double* d_relres, const int* d_status, int* h_iters,
                          double* h_relres, int* h_status) {
  const int nrhs = gcount * S->nsp;
  std::vector<int> it(nrhs), stt(nrhs);
  std::vector<double> rr(nrhs);
  DP_HIP(hipMemcpyAsync(it.data(), d_iters, nrhs * 4, hipMemcpyDeviceToHost, st));
  DP_HIP(hipMemcpyAsync(stt.data(), d_status, nrhs * 4, hipMemcpyDeviceToHost, st));
  DP_HIP(hipMemcpyAsync(rr.data(), d_relres, nrhs * 8, hipMemcpyDeviceToHost, st));
  DP_HIP(hipStreamSynchronize(st));
  for (int g = 0; g < gcount; ++g) {
    const int gg = g0 + g, f = gg / S->ncb, cb = gg % S->ncb;
    for (int c = 0; c < S->nsp; ++c) {
      const int col = cb * S->nsp + c;
      if (col >= S->ncols) continue;
      const int o = f * S->ncols + col, i = g * S->nsp + c;
      int code = stt[i];
      if (code < 0) code = DOGPUP_RHS_MAXIT;
      if (h_iters) h_iters[o] = it[i];
      if (h_relres) h_relres[o] = rr[i];
      if (h_status) h_status[o] = code;
    }
  }
  return DOGPUP_OK;
}

// Row permutation between the caller's node order and the solver's.
static int permute_block(dogpup_solver* S, hipStream_t st, const double* src, double* dst, int gc,
                         int dir) {
  const long long n = (long long)gc * S->nn * S->nsp;
  k_permute_rows<<<grid1(n), 256, 0, st>>>(src, dst, S->order.as<int>(), S->nn, S->nsp, gc, dir);
  DP_LAUNCH_CHECK();
  return DOGPUP_OK;
}

template <int NSP>
static int solve_c128(dogpup_solver* S, hipStream_t st, const double* b, double* x, double tol,
                      int maxit, int* h_iters, double* h_relres, int* h_status) {
  const size_t per_group = (size_t)S->nn * NSP * 2 * sizeof(double);
  for (int g0 = 0; g0 < S->ngroups; g0 += S->chunk) {
    const int gc = std::min(S->chunk, S->ngroups - g0);
    Chunk<double, NSP> C{S, st, gc, g0, carve_state(S->state, gc * NSP, gc), S->part1.as<double>(),
                         S->part2.as<double>()};
    Vecs<double> V = carve<double>(S->ws, gc * per_group, false, false);
    const size_t goff = (size_t)g0 * S->nn * NSP * 2;
    DP_TRY(permute_block(S, st, b + goff, S->bperm.as<double>(), gc, 0));
    V.x = S->xperm.as<double>();
    const double* bc = S->bperm.as<double>();
    k_fill<<<div_up(gc * NSP, 256), 256, 0, st>>>(C.bs.tol, tol, gc * NSP);
    DP_LAUNCH_CHECK();
    int its = 0;
    DP_TRY(C.run_fused(V, bc, maxit, nullptr, &its));
    S->stats[0] = std::max(S->stats[0], (double)its);
    // true residual for reporting
    k_residual64<NSP><<<dim3(S->nbx, gc), 256, 0, st>>>(
        S->rA, S->cA, S->a64pre.as<double>(), S->a64pim.as<double>(), S->fscale.as<double>(), bc, V.x,
        nullptr, nullptr, nullptr, nullptr, nullptr, 0, C.p1, C.p2, S->nn, S->ncb, g0, S->nbx, S->tile_rows);
    DP_LAUNCH_CHECK();
    // relres = |b - Ax| / |b|: reuse k_rf_check's arithmetic through a small pass
    std::vector<double> part((size_t)gc * S->nbx * NSP * 2);
    DP_HIP(hipMemcpyAsync(part.data(), C.p1, part.size() * 8, hipMemcpyDeviceToHost, st));
    std::vector<double> bn(gc * NSP);
    DP_HIP(hipMemcpyAsync(bn.data(), C.bs.bnorm, bn.size() * 8, hipMemcpyDeviceToHost, st));
    DP_TRY(permute_block(S, st, V.x, x + goff, gc, 1));
    DP_TRY(export_results(S, st, g0, gc, C.bs.iters, C.bs.relres, C.bs.status, h_iters, h_relres,
                          h_status));
    for (int g = 0; g < gc; ++g) {
      const int gg = g0 + g, f = gg / S->ncb, cb = gg % S->ncb;
      for (int c = 0; c < NSP; ++c) {
        const int col = cb * NSP + c;
        if (col >= S->ncols) continue;
        double s2 = 0.0;
        for (int bx = 0; bx < S->nbx; ++bx) s2 += part[2 * (((size_t)g * S->nbx + bx) * NSP + c)];
        const int i = g * NSP + c;
        if (h_relres) h_relres[f * S->ncols + col] = bn[i] > 0 ? std::sqrt(s2) / bn[i] : 0.0;
      }
    }
    double sum_it = 0;
    if (h_iters)
      for (int g = 0; g < gc; ++g) {
        const int gg = g0 + g, f = gg / S->ncb, cb = gg % S->ncb;
        int mx = 0;
        for (int c = 0; c < NSP; ++c) {
          const int col = cb * NSP + c;
          if (col < S->ncols) mx = std::max(mx, h_iters[f * S->ncols + col]);
        }
        sum_it += mx;
      }
    S->stats[1] += sum_it;
  }
  return DOGPUP_OK;
}

template <int NSP>
static int solve_refine(dogpup_solver* S, hipStream_t st, const double* b, double* x, double tol,
                        int maxit, int* h_iters, double* h_relres, int* h_status) {
  const size_t per_group32 = (size_t)S->nn * NSP * 2 * sizeof(float);
  const char* env_tol = getenv("DOGPUP_INNER_TOL");
  const double tol_in = env_tol ? atof(env_tol) : 1e-4;
  for (int g0 = 0; g0 < S->ngroups; g0 += S->chunk) {
    const int gc = std::min(S->chunk, S->ngroups - g0);
    Chunk<float, NSP> C{S, st, gc, g0, carve_state(S->state, gc * NSP, gc), S->part1.as<double>(),
                        S->part2.as<double>()};
    RState R = carve_rstate(S->rstate, gc * NSP);
    Vecs<float> V = carve<float>(S->ws, gc * per_group32, true, false);
    const size_t goff = (size_t)g0 * S->nn * NSP * 2;
    DP_TRY(permute_block(S, st, b + goff, S->bperm.as<double>(), gc, 0));
    double* xc = S->xperm.as<double>();
    const double* bc = S->bperm.as<double>();
    DP_HIP(hipMemsetAsync(xc, 0, (size_t)gc * S->nn * NSP * 16, st));
    const double e64 = (double)S->nn * NSP;  // complex elements per group
    for (int sweep = 0;; ++sweep) {
      C.pbegin();
      k_residual64<NSP><<<dim3(S->nbx, gc), 256, 0, st>>>(
          S->rA, S->cA, S->a64pre.as<double>(), S->a64pim.as<double>(), S->fscale.as<double>(), bc,
          sweep == 0 ? nullptr : xc, V.r, S->method == DOGPUP_METHOD_COCG ? nullptr : V.rhat, V.p, V.x,
          sweep == 0 ? nullptr : R.status, 1, C.p1, C.p2, S->nn, S->ncb, g0, S->nbx, S->tile_rows);
      DP_LAUNCH_CHECK();
      C.pend(3, (sweep == 0 ? gc : S->active_groups) * e64 * (16 + 16 + 4 * 8) +
                    (double)S->nnzA * 20 + 4.0 * (S->nn + 1));
      DP_TRY(C.reset_count());
      k_rf_check<<<gc, kScalThreads, 0, st>>>(R, C.bs, C.p1, NSP, S->nbx, S->ncb, S->ncols, g0, sweep, tol,
                                    tol_in, maxit);
      DP_LAUNCH_CHECK();
      int active = 0;
      DP_TRY(C.wait_count(&active));
      if (active == 0) break;
      const int active_groups_at_sweep = S->active_groups;
      // inner state from the residual pass: bnorm = |r64|, rho = r32.r32
      k_sc_init<<<gc, kScalThreads, 0, st>>>(C.bs, C.p1, C.p2, R.status, NSP, S->nbx, S->ncb, S->ncols, g0);
      DP_LAUNCH_CHECK();
      int its = 0;
      DP_TRY(C.run_fused(V, nullptr, maxit, R.status, &its));
      const int sweep_groups = active_groups_at_sweep;
      C.pbegin();
      k_add_correction<NSP><<<dim3(grid1((long long)S->nn * NSP / 4), gc), 256, 0, st>>>(
          xc, V.x, C.bs.status, S->nn);
      DP_LAUNCH_CHECK();
      C.pend(3, sweep_groups * e64 * (32 + 8));
      k_rf_accum<<<gc, kScalThreads, 0, st>>>(R, C.bs, NSP);
      DP_LAUNCH_CHECK();
      S->stats[0] = std::max(S->stats[0], (double)its);
      S->stats[2] += 1;
      // active-group-iterations for bandwidth accounting (host side estimate)
      S->stats[3] += its;
      if (sweep > 64) break;
    }
    DP_TRY(permute_block(S, st, xc, x + goff, gc, 1));
    DP_TRY(export_results(S, st, g0, gc, R.iters, R.relres, R.status, h_iters, h_relres, h_status));
  }
  return DOGPUP_OK;
}

template <int NSP>
static int solve_plain(dogpup_solver* S, hipStream_t st, const double* b, double* x, double tol,
                       int maxit, int* h_iters, double* h_relres, int* h_status) {
  const size_t per_group = (size_t)S->nn * NSP * 2 * sizeof(double);
  for (int g0 = 0; g0 < S->ngroups; g0 += S->chunk) {
    const int gc = std::min(S->chunk, S->ngroups - g0);
    Chunk<double, NSP> C{S, st, gc, g0, carve_state(S->state, gc * NSP, gc), S->part1.as<double>(),
                         S->part2.as<double>()};
    Vecs<double> V = carve<double>(S->ws, gc * per_group, false, true);
    const size_t goff = (size_t)g0 * S->nn * NSP * 2;
    DP_TRY(permute_block(S, st, b + goff, S->bperm.as<double>(), gc, 0));
    V.x = S->xperm.as<double>();
    const double* bc = S->bperm.as<double>();
    unsigned long long* dmax = reinterpret_cast<unsigned long long*>(C.bs.counter);
    const dim3 vg = C.vgrid();
    const int nn = S->nn, nbx = S->nbx;
    auto maxerr = [&](int op, const double* pa, double* out) -> int {
      DP_HIP(hipMemsetAsync(dmax, 0, 8, st));
      k_plain_scal<<<gc, kScalThreads, 0, st>>>(C.bs, op, pa, nullptr, NSP, nbx, S->ncb, S->ncols, g0, dmax);
      DP_LAUNCH_CHECK();
      DP_HIP(hipMemcpyAsync(S->h_max, dmax, 8, hipMemcpyDeviceToHost, st));
      DP_HIP(hipStreamSynchronize(st));
      unsigned long long bits = *S->h_max;
      double e;
      std::memcpy(&e, &bits, 8);
      *out = e;
      return DOGPUP_OK;
    };
    // x = 0, r = b, rhat0 = r, p = r
    k_init<double, NSP><<<vg, 256, 0, st>>>(bc, V.x, V.r, V.rhat, V.p, C.p1, C.p2, nn, nbx, S->tile_rows, nullptr);
    DP_LAUNCH_CHECK();
    k_plain_scal<<<gc, kScalThreads, 0, st>>>(C.bs, 5, C.p1, C.p2, NSP, nbx, S->ncb, S->ncols, g0, dmax);
    DP_LAUNCH_CHECK();
    int it = 0;
    for (it = 1; it <= maxit; ++it) {
      DP_TRY(C.template spmvG<EPI_STORE>(false, V.p, V.u, nullptr, nullptr));   // h = G p
      DP_TRY(C.template spmvG<EPI_STORE>(true, V.u, V.h, nullptr, nullptr));    // h = G^T h
      DP_TRY(C.template spmvA<EPI_STORE>(V.h, V.v, nullptr, nullptr));          // v = A h
      k_dot<double, NSP><<<vg, 256, 0, st>>>(V.v, V.rhat, 0, C.p1, nn, nbx, S->tile_rows, nullptr);
      DP_LAUNCH_CHECK();
      k_plain_scal<<<gc, kScalThreads, 0, st>>>(C.bs, 0, C.p1, nullptr, NSP, nbx, S->ncb, S->ncols, g0, dmax);
      DP_LAUNCH_CHECK();
      k_axpby<double, NSP><<<vg, 256, 0, st>>>(V.u, V.x, V.h, C.bs.alpha, 0, nn, nbx, S->tile_rows, nullptr);  // hx = x + h/alph
      DP_LAUNCH_CHECK();
      k_axpby<double, NSP><<<vg, 256, 0, st>>>(V.s, V.r, V.v, C.bs.alpha, 1, nn, nbx, S->tile_rows, nullptr);  // s = r - v/alph
      DP_LAUNCH_CHECK();
      k_dot<double, NSP><<<vg, 256, 0, st>>>(V.s, nullptr, 1, C.p1, nn, nbx, S->tile_rows, nullptr);
      DP_LAUNCH_CHECK();
      double err = 0;
      DP_TRY(maxerr(1, C.p1, &err));
      if (err < tol) {
        DP_HIP(hipMemcpyAsync(V.x, V.u, (size_t)gc * per_group, hipMemcpyDeviceToDevice, st));
        break;
      }
      std::swap(V.h, V.u);  // h now holds x + h/alph
      DP_TRY(C.template spmvG<EPI_STORE>(false, V.s, V.u, nullptr, nullptr));   // Gs
      DP_TRY(C.template spmvG<EPI_STORE>(true, V.u, V.z, nullptr, nullptr));    // z
      DP_TRY(C.template spmvA<EPI_STORE>(V.z, V.t, nullptr, nullptr));          // t
      // Gtt = G t is materialised, as the reference does
      DP_TRY(C.template spmvG<EPI_STORE>(false, V.t, V.w, nullptr, nullptr));
      k_dot<double, NSP><<<vg, 256, 0, st>>>(V.w, V.u, 0, C.p1, nn, nbx, S->tile_rows, nullptr);
      DP_LAUNCH_CHECK();
      k_dot<double, NSP><<<vg, 256, 0, st>>>(V.w, V.w, 0, C.p2, nn, nbx, S->tile_rows, nullptr);
      DP_LAUNCH_CHECK();
      k_plain_scal<<<gc, kScalThreads, 0, st>>>(C.bs, 2, C.p1, C.p2, NSP, nbx, S->ncb, S->ncols, g0, dmax);
      DP_LAUNCH_CHECK();
      k_axpby<double, NSP><<<vg, 256, 0, st>>>(V.x, V.h, V.z, C.bs.omega, 2, nn, nbx, S->tile_rows, nullptr);  // x = h + om z
      DP_LAUNCH_CHECK();
      k_axpby<double, NSP><<<vg, 256, 0, st>>>(V.r, V.s, V.t, C.bs.omega, 3, nn, nbx, S->tile_rows, nullptr);  // r = s - om t
      DP_LAUNCH_CHECK();
      k_axpby<double, NSP><<<vg, 256, 0, st>>>(V.p, V.p, V.v, C.bs.omega, 3, nn, nbx, S->tile_rows, nullptr);  // p = p - om v
      DP_LAUNCH_CHECK();
      k_dot<double, NSP><<<vg, 256, 0, st>>>(V.r, nullptr, 1, C.p1, nn, nbx, S->tile_rows, nullptr);
      DP_LAUNCH_CHECK();
      DP_TRY(maxerr(3, C.p1, &err));
      if (err < tol) break;
      k_dot<double, NSP><<<vg, 256, 0, st>>>(V.rhat, V.r, 0, C.p1, nn, nbx, S->tile_rows, nullptr);
      DP_LAUNCH_CHECK();
      k_plain_scal<<<gc, kScalThreads, 0, st>>>(C.bs, 4, C.p1, nullptr, NSP, nbx, S->ncb, S->ncols, g0, dmax);
      DP_LAUNCH_CHECK();
      k_axpby<double, NSP><<<vg, 256, 0, st>>>(V.p, V.r, V.p, C.bs.beta, 2, nn, nbx, S->tile_rows, nullptr);  // p = r + beta p
      DP_LAUNCH_CHECK();
    }
    DP_TRY(permute_block(S, st, V.x, x + goff, gc, 1));
    S->stats[0] = std::max(S->stats[0], (double)std::min(it, maxit));
    S->stats[1] += (double)std::min(it, maxit) * gc;
    // status: converged if the lockstep loop met tol
    std::vector<int> stt(gc * NSP);
    std::vector<double> rr(gc * NSP), bn(gc * NSP);
    DP_HIP(hipMemcpyAsync(rr.data(), C.bs.relres, rr.size() * 8, hipMemcpyDeviceToHost, st));
    DP_HIP(hipMemcpyAsync(bn.data(), C.bs.bnorm, bn.size() * 8, hipMemcpyDeviceToHost, st));
    DP_HIP(hipStreamSynchronize(st));
    for (int g = 0; g < gc; ++g) {
      const int gg = g0 + g, f = gg / S->ncb, cb = gg % S->ncb;
      for (int c = 0; c < NSP; ++c) {
        const int col = cb * NSP + c;
        if (col >= S->ncols) continue;
        const int o = f * S->ncols + col, i = g * NSP + c;
        if (h_iters) h_iters[o] = std::min(it, maxit);
        if (h_relres) h_relres[o] = rr[i];
        if (h_status) h_status[o] = bn[i] == 0.0 ? DOGPUP_RHS_ZERO
                                    : (rr[i] < tol ? DOGPUP_RHS_CONVERGED : DOGPUP_RHS_MAXIT);
      }
    }
  }
  return DOGPUP_OK;
}

}  // namespace dogpup

namespace dogpup {

// Solver-layout transforms and point-source right-hand sides
__global__ void k_to_solver(const double* __restrict__ src, double* __restrict__ dst, int nn, int nf,
                            int ncols, int nsp, int ncb) {
  // dst (g, node, c) <- src(node, f, col) (MATLAB NN x NF x NS), zero padding
  const long long total = (long long)nf * ncb * nn * nsp;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(i % nsp);
    const long long rest = i / nsp;
    const int node = (int)(rest % nn);
    const int g = (int)(rest / nn);
    const int f = g / ncb, col = (g % ncb) * nsp + c;
    double re = 0.0, im = 0.0;
    if (col < ncols) {
      const long long s = node + (long long)nn * (f + (long long)nf * col);
      re = src[2 * s];
      im = src[2 * s + 1];
    }
    dst[2 * i] = re;
    dst[2 * i + 1] = im;
  }
}

__global__ void k_from_solver(const double* __restrict__ src, double* __restrict__ dst, int nn,
                              int nf, int ncols, int nsp, int ncb) {
  const long long total = (long long)nn * nf * ncols;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int node = (int)(i % nn);
    const long long rest = i / nn;
    const int f = (int)(rest % nf);
    const int col = (int)(rest / nf);
    const int g = f * ncb + col / nsp, c = col % nsp;
    const long long s = ((long long)g * nn + node) * nsp + c;
    dst[2 * i] = src[2 * s];
    dst[2 * i + 1] = src[2 * s + 1];
  }
}

__global__ void k_point_sources(double* __restrict__ b, int nn, int nf, int ncols, int nsp, int ncb,
                                const int* __restrict__ nodes, const double* __restrict__ w,
                                const double* __restrict__ spec) {
  // one thread per (freq, col, q)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf * ncols * 4) return;
  const int q = i % 4, col = (i / 4) % ncols, f = i / (4 * ncols);
  const int node = nodes[col * 4 + q];
  const double wq = w[col * 4 + q];
  if (node < 0 || node >= nn || wq == 0.0) return;
  const int g = f * ncb + col / nsp, c = col % nsp;
  const long long o = ((long long)g * nn + node) * nsp + c;
  atomicAdd(&b[2 * o], wq * spec[2 * f]);
  atomicAdd(&b[2 * o + 1], wq * spec[2 * f + 1]);
}

static int check_nsp(int nsp) { return nsp == 8 || nsp == 16 || nsp == 32 || nsp == 64; }

template <template <int> class F, typename... A>
static int by_nsp(int nsp, A... a) {
  switch (nsp) {
    case 8: return F<8>::run(a...);
    case 16: return F<16>::run(a...);
    case 32: return F<32>::run(a...);
    case 64: return F<64>::run(a...);
  }
  return fail(DOGPUP_ERR_ARG, "unsupported group width %d", nsp);
}

template <int NSP>
struct SolveC128 {
  static int run(dogpup_solver* S, hipStream_t st, const double* b, double* x, double tol, int maxit,
                 int* hi, double* hr, int* hs) {
    return solve_c128<NSP>(S, st, b, x, tol, maxit, hi, hr, hs);
  }
};
template <int NSP>
struct SolveRefine {
  static int run(dogpup_solver* S, hipStream_t st, const double* b, double* x, double tol, int maxit,
                 int* hi, double* hr, int* hs) {
    return solve_refine<NSP>(S, st, b, x, tol, maxit, hi, hr, hs);
  }
};
template <int NSP>
struct SolvePlain {
  static int run(dogpup_solver* S, hipStream_t st, const double* b, double* x, double tol, int maxit,
                 int* hi, double* hr, int* hs) {
    return solve_plain<NSP>(S, st, b, x, tol, maxit, hi, hr, hs);
  }
};

static double ws_budget_bytes() {
  const char* e = getenv("DOGPUP_WORKSPACE_GB");
  double gb = e ? atof(e) : 160.0;
  if (gb <= 0) gb = 160.0;
  return gb * 1073741824.0;
}

}  // namespace dogpup

extern "C" {

int dogpup_solver_create(dogpup_solver_t* out, int nn, int nfreq, int ncols, int nsp, int mode,
                         int nnzA, const int* d_rptrA, const int* d_cptrA, int nnzG,
                         const int* d_rptrG, const int* d_cptrG, const int* h_order) {
  DP_ARG(out != nullptr, "out is NULL");
  *out = nullptr;
  DP_ARG(nn > 0 && nfreq > 0 && ncols > 0 && nnzA > 0 && nnzG > 0, "bad dimensions");
  DP_ARG(check_nsp(nsp), "group width must be 8, 16, 32 or 64 (got %d)", nsp);
  DP_ARG(mode == DOGPUP_MODE_C128 || mode == DOGPUP_MODE_C64_REFINE || mode == DOGPUP_MODE_C128_PLAIN,
         "unknown mode %d", mode);
  DP_ARG(d_rptrA && d_cptrA && d_rptrG && d_cptrG, "NULL pattern");
  std::unique_ptr<dogpup_solver> S(new dogpup_solver());
  S->nn = nn; S->nf = nfreq; S->ncols = ncols; S->nsp = nsp; S->mode = mode;
  S->ncb = (ncols + nsp - 1) / nsp;
  S->ngroups = nfreq * S->ncb;
  S->nnzA = nnzA; S->nnzG = nnzG;
  S->rA0 = d_rptrA; S->cA0 = d_cptrA; S->rG0 = d_rptrG; S->cG0 = d_cptrG;
  // Row tile of the vector/SpMV kernels (a multiple of 64 rows); gridDim.x
  // is a multiple of 8 (one contiguous tile range per XCD).
  if (const char* e = getenv("DOGPUP_TILE")) S->tile_rows = atoi(e);
  if (S->tile_rows < 64 || S->tile_rows % 64) S->tile_rows = 512;
  if (const char* e = getenv("DOGPUP_SPMV_LDS_PAD")) S->lds_pad = atoi(e);
  S->nbx = (int)(div_up(div_up(nn, S->tile_rows), 8) * 8);
  DP_ARG(S->ngroups <= 65535, "too many groups (%d)", S->ngroups);

  // solver node order and the permuted patterns (host-side, once per handle)
  std::vector<int> order(nn), inv(nn, -1);
  for (int i = 0; i < nn; ++i) order[i] = h_order ? h_order[i] : i;
  for (int i = 0; i < nn; ++i) {
    DP_ARG(order[i] >= 0 && order[i] < nn && inv[order[i]] < 0, "node order is not a permutation");
    inv[order[i]] = i;
  }
  std::vector<int> rA(nn + 1), cA(nnzA), rG(nn + 1), cG(nnzG);
  DP_HIP(hipMemcpy(rA.data(), d_rptrA, (nn + 1) * 4, hipMemcpyDeviceToHost));
  DP_HIP(hipMemcpy(cA.data(), d_cptrA, (size_t)nnzA * 4, hipMemcpyDeviceToHost));
  DP_HIP(hipMemcpy(rG.data(), d_rptrG, (nn + 1) * 4, hipMemcpyDeviceToHost));
  DP_HIP(hipMemcpy(cG.data(), d_cptrG, (size_t)nnzG * 4, hipMemcpyDeviceToHost));
  DP_ARG(rA[0] == 0 && rA[nn] == nnzA && rG[0] == 0 && rG[nn] == nnzG, "CSR row pointers inconsistent with nnz");
  for (int i = 0; i < nn; ++i) DP_ARG(rA[i] <= rA[i + 1] && rG[i] <= rG[i + 1], "CSR row pointers decrease");
  for (int v : cA) DP_ARG(v >= 0 && v < nn, "A column index out of range");
  for (int v : cG) DP_ARG(v >= 0 && v < nn, "G column index out of range");
  auto permute = [&](const std::vector<int>& rp, const std::vector<int>& cp, std::vector<int>& rq,
                     std::vector<int>& cq, std::vector<int>& map) {
    rq.assign(nn + 1, 0);
    cq.resize(cp.size());
    map.resize(cp.size());
    for (int i = 0; i < nn; ++i) rq[i + 1] = rq[i] + (rp[order[i] + 1] - rp[order[i]]);
    std::vector<std::pair<int, int>> row;
    for (int i = 0; i < nn; ++i) {
      const int o = order[i];
      row.clear();
      for (int p = rp[o]; p < rp[o + 1]; ++p) row.push_back({inv[cp[p]], p});
      std::sort(row.begin(), row.end());
      for (size_t k = 0; k < row.size(); ++k) {
        cq[rq[i] + k] = row[k].first;
        map[rq[i] + k] = row[k].second;
      }
    }
  };
  std::vector<int> rAp, cAp, amap, rGp, cGp, gmap;
  permute(rA, cA, rAp, cAp, amap);
  permute(rG, cG, rGp, cGp, gmap);
  // G^T in solver order: counting sort of the permuted G by column
  std::vector<int> rGT(nn + 1, 0), cGT(nnzG), gtmap(nnzG);
  for (int v : cGp) rGT[v + 1]++;
  for (int i = 0; i < nn; ++i) rGT[i + 1] += rGT[i];
  {
    std::vector<int> fill(rGT.begin(), rGT.end() - 1);
    for (int i = 0; i < nn; ++i)
      for (int p = rGp[i]; p < rGp[i + 1]; ++p) {
        const int q = fill[cGp[p]]++;
        cGT[q] = i;
        gtmap[q] = gmap[p];
      }
  }
  auto up = [&](dogpup::Buf& b, const std::vector<int>& v) -> int {
    DP_TRY(b.ensure(v.size() * 4));
    DP_HIP(hipMemcpy(b.p, v.data(), v.size() * 4, hipMemcpyHostToDevice));
    return DOGPUP_OK;
  };
  DP_TRY(up(S->pAr, rAp)); DP_TRY(up(S->pAc, cAp)); DP_TRY(up(S->amap, amap));
  DP_TRY(up(S->pGr, rGp)); DP_TRY(up(S->pGc, cGp)); DP_TRY(up(S->gmap, gmap));
  DP_TRY(up(S->pGTr, rGT)); DP_TRY(up(S->pGTc, cGT)); DP_TRY(up(S->gtmap, gtmap));
  DP_TRY(up(S->order, order));
  if (const char* e = getenv("DOGPUP_SPMV_UA")) S->unroll_a = atoi(e);
  if (const char* e = getenv("DOGPUP_SPMV_UG")) S->unroll_g = atoi(e);
  S->rA = S->pAr.as<int>(); S->cA = S->pAc.as<int>();
  S->rG = S->pGr.as<int>(); S->cG = S->pGc.as<int>();
  S->rGT = S->pGTr.as<int>(); S->cGT = S->pGTc.as<int>();

  const size_t csize = mode == DOGPUP_MODE_C64_REFINE ? 8 : 16;
  const size_t per_group = align256((size_t)nn * nsp * csize);
  const size_t per_group64 = align256((size_t)nn * nsp * 16);
  const int nvec = nvec_for(mode);
  // vector workspace + permuted b and x (complex128) per group
  long long chunk = (long long)(ws_budget_bytes() / ((double)per_group * nvec + 2.0 * per_group64));
  if (chunk < 1) chunk = 1;
  if (chunk > S->ngroups) chunk = S->ngroups;
  S->chunk = (int)chunk;
  DP_TRY(S->ws.ensure((size_t)nvec * align256(per_group * S->chunk)));
  DP_TRY(S->bperm.ensure(per_group64 * S->chunk));
  DP_TRY(S->xperm.ensure(per_group64 * S->chunk));
  const size_t pbytes = (size_t)S->chunk * S->nbx * nsp * 16;
  DP_TRY(S->part1.ensure(pbytes));
  DP_TRY(S->part2.ensure(pbytes));
  DP_TRY(S->state.ensure(state_bytes(S->chunk * nsp, S->chunk)));
  DP_TRY(S->rstate.ensure(5 * align256((size_t)S->chunk * nsp * 8)));
  DP_TRY(S->fscale.ensure((size_t)nfreq * 8));
  DP_HIP(hipHostMalloc((void**)&S->h_counter, 16, hipHostMallocDefault));
  DP_HIP(hipHostMalloc((void**)&S->h_max, 16, hipHostMallocDefault));
  DP_HIP(hipStreamCreateWithFlags(&S->own, hipStreamNonBlocking));
  DP_HIP(hipEventCreateWithFlags(&S->ev_in, hipEventDisableTiming));
  DP_HIP(hipEventCreateWithFlags(&S->ev_out, hipEventDisableTiming));
  if (const char* e = getenv("DOGPUP_GRAPH")) S->use_graph = atoi(e) != 0;
  if (const char* e = getenv("DOGPUP_GRAPH_K")) S->graph_k = std::max(1, atoi(e));
  *out = S.release();
  return DOGPUP_OK;
}

int dogpup_solver_destroy(dogpup_solver_t s) {
  if (s) {
    (void)hipDeviceSynchronize();
    delete s;
  }
  return DOGPUP_OK;
}

int dogpup_solver_set_matrix(dogpup_solver_t S, const double* d_valA_re, const double* d_valA_im,
                             const double* d_fscale, void* stream) {
  DP_ARG(S && d_valA_re && d_valA_im && d_fscale, "NULL argument");
  hipStream_t st = as_stream(stream);
  S->clear_graphs();
  const size_t nA = S->nnzA;
  DP_TRY(S->a64re.ensure(nA * 8));
  DP_TRY(S->a64im.ensure(nA * 8));
  DP_HIP(hipMemcpyAsync(S->a64re.p, d_valA_re, nA * 8, hipMemcpyDeviceToDevice, st));
  DP_HIP(hipMemcpyAsync(S->a64im.p, d_valA_im, nA * 8, hipMemcpyDeviceToDevice, st));
  DP_HIP(hipMemcpyAsync(S->fscale.p, d_fscale, (size_t)S->nf * 8, hipMemcpyDeviceToDevice, st));
  // FSAI factors in the reference node order (pattern = the caller's tril(A))
  const size_t nG = (size_t)S->nnzG * S->nf;
  DP_TRY(S->g64.ensure(nG * 16));
  FsaiOut fo{S->g64.as<double>(), nullptr, S->nnzG, FSAI_OUT_SOLVER};
  DP_TRY(fsai_launch(S->a64re.as<double>(), S->a64im.as<double>(), 0, S->rA0, S->cA0, S->rG0, S->cG0,
                     S->nn, S->nf, S->fscale.as<double>(), fo, st));
  // values in solver node order
  DP_TRY(S->a64pre.ensure(nA * 8));
  DP_TRY(S->a64pim.ensure(nA * 8));
  k_gather_real<double><<<grid1(nA), 256, 0, st>>>(S->a64re.as<double>(), S->amap.as<int>(), S->a64pre.as<double>(), nA);
  k_gather_real<double><<<grid1(nA), 256, 0, st>>>(S->a64im.as<double>(), S->amap.as<int>(), S->a64pim.as<double>(), nA);
  DP_LAUNCH_CHECK();
  if (S->mode == DOGPUP_MODE_C64_REFINE) {
    DP_TRY(S->a32re.ensure(nA * 4));
    DP_TRY(S->a32im.ensure(nA * 4));
    k_gather_real<float><<<grid1(nA), 256, 0, st>>>(S->a64re.as<double>(), S->amap.as<int>(), S->a32re.as<float>(), nA);
    k_gather_real<float><<<grid1(nA), 256, 0, st>>>(S->a64im.as<double>(), S->amap.as<int>(), S->a32im.as<float>(), nA);
    DP_TRY(S->g32.ensure(nG * 8));
    DP_TRY(S->gt32.ensure(nG * 8));
    k_gather_cplx<float><<<grid1(nG), 256, 0, st>>>(S->g64.as<double>(), S->gmap.as<int>(), S->g32.as<float>(), S->nnzG, S->nf);
    k_gather_cplx<float><<<grid1(nG), 256, 0, st>>>(S->g64.as<double>(), S->gtmap.as<int>(), S->gt32.as<float>(), S->nnzG, S->nf);
    DP_LAUNCH_CHECK();
  } else {
    DP_TRY(S->gp64.ensure(nG * 16));
    DP_TRY(S->gtp64.ensure(nG * 16));
    k_gather_cplx<double><<<grid1(nG), 256, 0, st>>>(S->g64.as<double>(), S->gmap.as<int>(), S->gp64.as<double>(), S->nnzG, S->nf);
    k_gather_cplx<double><<<grid1(nG), 256, 0, st>>>(S->g64.as<double>(), S->gtmap.as<int>(), S->gtp64.as<double>(), S->nnzG, S->nf);
    DP_LAUNCH_CHECK();
  }
  S->have_matrix = true;
  return DOGPUP_OK;
}

int dogpup_solver_solve(dogpup_solver_t S, const double* d_b, double* d_x, double tol, int maxit,
                        int* h_iters, double* h_relres, int* h_status, void* stream) {
  DP_ARG(S && d_b && d_x, "NULL argument");
  DP_ARG(S->have_matrix, "dogpup_solver_set_matrix has not been called");
  DP_ARG(tol > 0 && maxit > 0, "tol and maxit must be positive");
  // run on the handle's own stream (capturable), ordered after the caller's
  hipStream_t caller = as_stream(stream);
  DP_HIP(hipEventRecord(S->ev_in, caller));
  DP_HIP(hipStreamWaitEvent(S->own, S->ev_in, 0));
  hipStream_t st = S->own;
  for (double& v : S->stats) v = 0.0;
  int rc;
  switch (S->mode) {
    case DOGPUP_MODE_C128:
      rc = by_nsp<SolveC128>(S->nsp, S, st, d_b, d_x, tol, maxit, h_iters, h_relres, h_status);
      break;
    case DOGPUP_MODE_C64_REFINE:
      rc = by_nsp<SolveRefine>(S->nsp, S, st, d_b, d_x, tol, maxit, h_iters, h_relres, h_status);
      break;
    default:
      rc = by_nsp<SolvePlain>(S->nsp, S, st, d_b, d_x, tol, maxit, h_iters, h_relres, h_status);
  }
  if (rc != DOGPUP_OK) return rc;
  DP_HIP(hipEventRecord(S->ev_out, st));
  DP_HIP(hipStreamWaitEvent(caller, S->ev_out, 0));
  DP_HIP(hipStreamSynchronize(st));
  for (int c = 0; c < 4; ++c) S->prof_ms[c] = S->prof_bytes[c] = S->prof_n[c] = 0.0;
  for (auto& r : S->recs) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      S->prof_ms[r.cls] += ms;
      S->prof_bytes[r.cls] += r.bytes;
      S->prof_n[r.cls] += 1;
    }
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  S->recs.clear();
  return DOGPUP_OK;
}

int dogpup_solver_set_method(dogpup_solver_t S, int method) {
  DP_ARG(S != nullptr, "NULL solver");
  DP_ARG(method == DOGPUP_METHOD_BICGSTAB || method == DOGPUP_METHOD_COCG, "unknown method %d", method);
  DP_ARG(method == DOGPUP_METHOD_BICGSTAB || S->mode != DOGPUP_MODE_C128_PLAIN,
         "the reference-structure form is BiCGSTAB only");
  if (method != S->method) S->clear_graphs();
  S->method = method;
  return DOGPUP_OK;
}

int dogpup_solver_set_profile(dogpup_solver_t S, int on) {
  DP_ARG(S != nullptr, "NULL solver");
  S->profile = on != 0;
  return DOGPUP_OK;
}

int dogpup_solver_profile(dogpup_solver_t S, double* h_out, int n) {
  DP_ARG(S && h_out, "NULL argument");
  for (int c = 0; c < 4; ++c) {
    if (3 * c + 2 < n) {
      h_out[3 * c] = S->prof_ms[c];
      h_out[3 * c + 1] = S->prof_bytes[c];
      h_out[3 * c + 2] = S->prof_n[c];
    }
  }
  return DOGPUP_OK;
}

int dogpup_solver_stats(dogpup_solver_t S, double* h_out, int n) {
  DP_ARG(S && h_out, "NULL argument");
  for (int i = 0; i < n && i < 4; ++i) h_out[i] = S->stats[i];
  return DOGPUP_OK;
}

int dogpup_solver_get_fsai(dogpup_solver_t S, double* d_out, void* stream) {
  DP_ARG(S && d_out, "NULL argument");
  DP_ARG(S->have_matrix, "no matrix set");
  DP_HIP(hipMemcpyAsync(d_out, S->g64.p, (size_t)S->nnzG * S->nf * 16, hipMemcpyDeviceToDevice,
                        as_stream(stream)));
  return DOGPUP_OK;
}

int dogpup_layout_to_solver(const double* d_src, double* d_dst, int nn, int nfreq, int ncols,
                            int nsp, void* stream) {
  DP_ARG(d_src && d_dst && nn > 0 && nfreq > 0 && ncols > 0 && check_nsp(nsp), "bad arguments");
  const int ncb = (ncols + nsp - 1) / nsp;
  k_to_solver<<<grid1((long long)nfreq * ncb * nn * nsp), 256, 0, as_stream(stream)>>>(
      d_src, d_dst, nn, nfreq, ncols, nsp, ncb);
  DP_LAUNCH_CHECK();
  return DOGPUP_OK;
}

int dogpup_layout_from_solver(const double* d_src, double* d_dst, int nn, int nfreq, int ncols,
                              int nsp, void* stream) {
  DP_ARG(d_src && d_dst && nn > 0 && nfreq > 0 && ncols > 0 && check_nsp(nsp), "bad arguments");
  const int ncb = (ncols + nsp - 1) / nsp;
  k_from_solver<<<grid1((long long)nn * nfreq * ncols), 256, 0, as_stream(stream)>>>(
      d_src, d_dst, nn, nfreq, ncols, nsp, ncb);
  DP_LAUNCH_CHECK();
  return DOGPUP_OK;
}

int dogpup_point_sources(double* d_b, int nn, int nfreq, int ncols, int nsp, const int* d_nodes,
                         const double* d_w, const double* d_spectrum, void* stream) {
  DP_ARG(d_b && d_nodes && d_w && d_spectrum && nn > 0 && nfreq > 0 && ncols > 0 && check_nsp(nsp),
         "bad arguments");
  const int ncb = (ncols + nsp - 1) / nsp;
  hipStream_t st = as_stream(stream);
  DP_HIP(hipMemsetAsync(d_b, 0, (size_t)nfreq * ncb * nn * nsp * 16, st));
  const int n = nfreq * ncols * 4;
  k_point_sources<<<div_up(n, 256), 256, 0, st>>>(d_b, nn, nfreq, ncols, nsp, ncb, d_nodes, d_w,
                                                   d_spectrum);
  DP_LAUNCH_CHECK();
  return DOGPUP_OK;
}

}  // extern "C"
