// Batched FSAI-preconditioned Krylov solves over every (mesh, frequency,
// column) RHS.
//
// Algorithm (behaviour of DOGPUP/dMesh.m:1247-1348, gpuBicstab_FSAIP):
// right-preconditioned BiCGSTAB on A_k x = b with M^-1 = G_k^T G_k, x0 = 0,
// r_hat0 = r0, unconjugated bilinear forms, omega = (Gt . Gs)/(Gt . Gt)
// weighted by G, convergence on ||s||/||b|| or ||r||/||b||.  The fused forms
// can run COCG instead (CG with the unconjugated form, for the complex
// symmetric A_k and symmetric G_k^T G_k): same x, three operator passes per
// iteration instead of seven.
//
// Three forms share the CSR kernels:
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
//
// Batches of meshes.  A handle holds one or more independent meshes whose
// CSR patterns the caller concatenates into one block-diagonal pattern (mesh
// i owns rows [row0_i, row0_{i+1})).  A group is one (mesh, frequency, block
// of NSP columns); its vectors cover only its mesh's rows.  Every kernel
// launch walks one flat list of work items (group, row tile) over all groups
// of all meshes, so meshes of very different sizes fill one launch together.
#include <cmath>
#include <cstring>
#include <cstdint>
#include <cstdlib>
#include <algorithm>
#include <map>
#include <memory>
#include <type_traits>
#include <vector>

#include "fsai.hpp"
#include "kernels.hpp"
#include "rb_weights.hpp"

namespace dogpup {

enum { MAT_A = 0, MAT_G = 1 };
// SpMV epilogues: EPI_STORE y; EPI_STORE_DOT y and w.y; EPI_DOT2 only w.y
// and y.y (y not stored); EPI_STORE_SELF y and y.y; EPI_AXPBY y <- y_new +
// coef*y_old (in place, one complex coefficient per column); EPI_AXPBY_X
// also xacc += coef2*y_old.  All dots are unconjugated.
// EPI_AXPY_SELF y <- y_old - coef*(M x) (in place) with y.y and |y|^2;
// EPI_STORE_SELF2 y with y.y and |y|^2.
enum { EPI_STORE = 0, EPI_STORE_DOT = 1, EPI_DOT2 = 2, EPI_STORE_SELF = 3, EPI_AXPBY = 4, EPI_AXPBY_X = 5,
       EPI_AXPY_SELF = 6, EPI_STORE_SELF2 = 7 };
enum { ST_ACTIVE = -1, ST_SKIP = -2 };

// Work list.  Workgroups are dispatched in order, round-robin over the 8
// XCDs, so block b runs on XCD b % 8.  Each group's items are padded to a
// multiple of 8 and ordered so that XCD x sweeps the contiguous tile range
// [x*n/8, (x+1)*n/8) of the group (a tile's stencil neighbours then sit in
// the same XCD's L2), and the groups follow one another: all XCDs work on
// the same group at a time, whose vectors stay in the Infinity Cache.
// Padding items have tile -1 and do nothing.
struct Work {
  const int* wgrp;        // work item -> group
  const int* wtile;       // work item -> row tile within its group (-1: padding)
  const int* wpart;       // work item -> slot of its partial sums (dense over real items)
  const int* grow0;       // group -> first row of its mesh
  const int* grows;       // group -> rows of its mesh
  const long long* gvb;   // group -> node offset of its vector block
  const double* gfsc;     // group -> multiplier of Im A (frequency scale)
  const long long* ggv;   // group -> G value offset: entry p of the pattern is at ggv + p
  const int* gwbeg;       // group -> first partial-sum slot
  const int* gwcnt;       // group -> partial-sum slots (its row tiles)
  const int* gcols;       // group -> columns holding a RHS (the rest is padding)
  int w0, nw;             // this launch's work items [w0, w0 + nw), nw a multiple of 8
  int pw0;                // partial-sum slot of its first real item
  int g0;                 // this launch's first group
  long long vb0;          // node offset of that group's vector block
  int tile_rows;
};

// The block's work item: global group g, chunk-local group gl and item wl,
// its rows [row0, row1) and vb, the chunk-local node index of row 0 of its
// mesh (chunk-local node index of row r = vb + r).
struct Item {
  int g, gl, wl, row0, row1;
  long long vb;
};

__device__ inline bool work_item(const Work& W, Item& it) {
  if ((int)blockIdx.x >= W.nw) return false;
  const int w = W.w0 + blockIdx.x;
  const int t = W.wtile[w];
  if (t < 0) return false;
  const int g = W.wgrp[w];
  const int r0 = W.grow0[g];
  it.g = g;
  it.gl = g - W.g0;
  it.wl = W.wpart[w] - W.pw0;
  it.row0 = r0 + t * W.tile_rows;
  it.row1 = min(it.row0 + W.tile_rows, r0 + W.grows[g]);
  it.vb = W.gvb[g] - W.vb0 - r0;
  return true;
}

// ------------------------------------------------------------------ SpMV ---
// y = M x for every work item; MAT_A: M = v0 + i*gfsc[g]*v1 (planar, shared
// by all frequencies of a mesh); MAT_G: M = complex values per (mesh,
// frequency), entry p at v0[ggv[g] + p].
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
// loads unconditional.  A chunk with more than kChunkCap nonzeros reads the
// CSR arrays from global memory instead.
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
// (M x)(row) over this lane's CPL columns, handed to epi(row, ar, ai).  xg
// points at this lane's first column of x; row r of the mesh is at node vb + r.
// srp/scol/sval are the caller's LDS chunk buffers.  All threads of the block
// call it (it synchronises).
template <typename T, int NSP, int MAT, int U, typename Epi>
__device__ inline void spmv_rows(const int* __restrict__ rptr, const int* __restrict__ cptr,
                                 const T* __restrict__ v0, const T* __restrict__ v1, T fk,
                                 const T* __restrict__ xg, long long vb, int tile0, int tile1, int* srp,
                                 int* scol, T* sval, Epi&& epi) {
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
        for (int q = 0; q < U; ++q) Pack<T>::load(xg + (vb + cc[q]) * (NSP * 2), xr[q], xi[q]);
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

// Epilogue of the solver SpMV for one row: stores and fused dot products
// (see the EPI_* list above), shared by k_spmv and k_spmv_halo.
template <typename T, int NSP, int EPI>
struct SpmvEpi {
  static constexpr int CPL = Geo<T, NSP>::CPL;
  static constexpr bool AXPBY = EPI == EPI_AXPBY || EPI == EPI_AXPBY_X;
  T* y;
  const T* w;
  T* xacc;
  long long vb;
  int c0;
  T cr[CPL], ci[CPL], c2r[CPL], c2i[CPL];
  Acc<CPL> d1, d2;

  __device__ SpmvEpi(T* y_, const T* w_, T* xacc_, const double* coef, const double* coef2, long long vb_,
                     int c0_, long long rhs0)
      : y(y_), w(w_), xacc(xacc_), vb(vb_), c0(c0_) {
    d1.zero();
    d2.zero();
    if (AXPBY || EPI == EPI_AXPY_SELF) load_coef<T, CPL>(coef, rhs0, cr, ci);
    if (EPI == EPI_AXPBY_X) load_coef<T, CPL>(coef2, rhs0, c2r, c2i);
  }
  __device__ inline void operator()(int row, T* ar, T* ai) {
    const long long off = ((vb + row) * NSP + c0) * 2;
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
    if (EPI == EPI_AXPY_SELF) {
      T pr[CPL], pi[CPL];
      Pack<T>::load(y + off, pr, pi);
#pragma unroll
      for (int e = 0; e < CPL; ++e) {
        const T tr = ar[e], ti = ai[e];
        ar[e] = pr[e] - (cr[e] * tr - ci[e] * ti);
        ai[e] = pi[e] - (cr[e] * ti + ci[e] * tr);
      }
    }
    if (EPI != EPI_DOT2) Pack<T>::store(y + off, ar, ai);
    if (EPI == EPI_STORE_SELF || EPI == EPI_AXPY_SELF || EPI == EPI_STORE_SELF2) {
#pragma unroll
      for (int e = 0; e < CPL; ++e) {
        d1.re[e] += (double)ar[e] * ar[e] - (double)ai[e] * ai[e];
        d1.im[e] += 2.0 * (double)ar[e] * ai[e];
        if (EPI != EPI_STORE_SELF) d2.re[e] += (double)ar[e] * ar[e] + (double)ai[e] * ai[e];
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
  }
  // operands the epilogue reads (y_old or w, and xacc) for row `row`, as
  // 16-byte packs: the halo kernel loads them ahead of its compute phase
  static constexpr int NOPS = (EPI == EPI_AXPBY_X) ? 2
                              : (AXPBY || EPI == EPI_AXPY_SELF || EPI == EPI_STORE_DOT || EPI == EPI_DOT2) ? 1
                                                                                                      : 0;
  template <typename V>
  __device__ inline void load_ops(int row, V& o1, V& o2) const {
    const long long off = ((vb + row) * NSP + c0) * 2;
    if (NOPS >= 1) o1 = *reinterpret_cast<const V*>((EPI == EPI_STORE_DOT || EPI == EPI_DOT2 ? w : y) + off);
    if (NOPS >= 2) o2 = *reinterpret_cast<const V*>(xacc + off);
  }
  // operator() with the operands already loaded (CPL = 2, complex64)
  template <typename V>
  __device__ inline void apply(int row, T* ar, T* ai, const V& o1, const V& o2) {
    static_assert(CPL == 2 && sizeof(T) == 4, "complex64 packs");
    const long long off = ((vb + row) * NSP + c0) * 2;
    const T pr[2] = {o1[0], o1[2]}, pi[2] = {o1[1], o1[3]};
    if (AXPBY) {
      if (EPI == EPI_AXPBY_X) {
        T xr[2] = {o2[0], o2[2]}, xi[2] = {o2[1], o2[3]};
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
    if (EPI == EPI_AXPY_SELF) {
#pragma unroll
      for (int e = 0; e < CPL; ++e) {
        const T tr = ar[e], ti = ai[e];
        ar[e] = pr[e] - (cr[e] * tr - ci[e] * ti);
        ai[e] = pi[e] - (cr[e] * ti + ci[e] * tr);
      }
    }
    if (EPI != EPI_DOT2) Pack<T>::store(y + off, ar, ai);
    if (EPI == EPI_STORE_SELF || EPI == EPI_AXPY_SELF || EPI == EPI_STORE_SELF2) {
#pragma unroll
      for (int e = 0; e < CPL; ++e) {
        d1.re[e] += (double)ar[e] * ar[e] - (double)ai[e] * ai[e];
        d1.im[e] += 2.0 * (double)ar[e] * ai[e];
        if (EPI != EPI_STORE_SELF) d2.re[e] += (double)ar[e] * ar[e] + (double)ai[e] * ai[e];
      }
    }
    if (EPI == EPI_STORE_DOT || EPI == EPI_DOT2) {
#pragma unroll
      for (int e = 0; e < CPL; ++e) {
        d1.re[e] += (double)pr[e] * ar[e] - (double)pi[e] * ai[e];
        d1.im[e] += (double)pr[e] * ai[e] + (double)pi[e] * ar[e];
        if (EPI == EPI_DOT2) {
          d2.re[e] += (double)ar[e] * ar[e] - (double)ai[e] * ai[e];
          d2.im[e] += 2.0 * (double)ar[e] * ai[e];
        }
      }
    }
  }
  // the block's partial sums (all threads call it)
  __device__ inline void finish(double* sh, double* part1, double* part2, int wl) {
    if (EPI != EPI_STORE && !AXPBY) {
      block_reduce_store<T, NSP>(d1, sh, part1, wl);
      if (EPI == EPI_DOT2 || EPI == EPI_AXPY_SELF || EPI == EPI_STORE_SELF2) {
        __syncthreads();
        block_reduce_store<T, NSP>(d2, sh, part2, wl);
      }
    }
  }
};

// 6 waves per SIMD (at most 80 VGPRs) measured best for U = 8.
template <typename T, int NSP, int MAT, int EPI, int U>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(U <= 8 ? 6 : 4))) void k_spmv(
    const int* __restrict__ rptr, const int* __restrict__ cptr, const T* __restrict__ v0,
    const T* __restrict__ v1, const T* __restrict__ x, T* __restrict__ y, const T* __restrict__ w,
    const double* __restrict__ coef, const double* __restrict__ coef2, T* __restrict__ xacc,
    double* __restrict__ part1, double* __restrict__ part2, Work W, const int* __restrict__ grp_active) {
  using G = Geo<T, NSP>;
  constexpr int CPL = G::CPL;
  Item it;
  if (!work_item(W, it)) return;
  if (grp_active && !grp_active[it.gl]) return;
  __shared__ double sh[2 * 4 * NSP];
  __shared__ int srp[kChunkRows + 1];
  __shared__ int scol[kChunkCap];
  __shared__ T sval[2 * kChunkCap];
  const int lane = threadIdx.x & 63;
  const int c0 = (lane % G::LPR) * CPL;
  T fk = T(0);
  const T* gv = v0;
  if (MAT == MAT_A) fk = (T)W.gfsc[it.g];
  else gv = v0 + 2 * W.ggv[it.g];
  SpmvEpi<T, NSP, EPI> epi(y, w, xacc, coef, coef2, it.vb, c0, (long long)it.gl * NSP + c0);
  spmv_rows<T, NSP, MAT, U>(rptr, cptr, gv, v1, fk, x + c0 * 2, it.vb, it.row0, it.row1, srp, scol, sval, epi);
  epi.finish(sh, part1, part2, it.wl);
}

// x-halo form of k_spmv for complex64 (NSP <= 32).  For every 64-row chunk
// of a mesh the handle holds the sorted distinct columns its rows reference
// (the chunk's halo: about 184 rows for A and 124 for G or G^T on the
// 15-point Kuhn stencil, against 934 and 499 gathers) and, per nonzero, the
// column's slot in that list.  Each chunk's halo rows of x are copied into
// LDS once and every row product reads x from there, so L2 requests fall by
// the halo ratio (5.1x for A, 4.0x for G).
//
// The LDS holds one chunk (kHaloCap rows of NSP complex64 columns: two
// blocks per CU at NSP = 32), so latency is hidden by software pipelining
// over the chunks of the block's tile instead of by occupancy: while chunk c
// is computed from LDS, chunk c+1's x rows and matrix slice are in flight to
// registers and chunk c+2's halo list is being fetched.  Every load comes
// from an index clamped into range (a clamped lane writes the same value to
// the same slot as the lane that owns it), so no load sits behind a branch.
constexpr int kHaloCap = 240;
constexpr int kHaloTileRows = 1024;
// native 4-float vector: copies of it stay in registers (HIP's float4 struct
// is copied with memcpy, which kept the staged rows in scratch memory)
typedef float f4 __attribute__((ext_vector_type(4)));

// a += m * v for the two complex values of a 16-byte x piece, v = (re0,
// im0, re1, im1): two fused multiply-adds per real output.
__device__ inline void cmac2(const f4& v, float mr, float mi, float* ar, float* ai) {
  ar[0] = fmaf(mr, v[0], ar[0]);
  ar[0] = fmaf(-mi, v[1], ar[0]);
  ai[0] = fmaf(mr, v[1], ai[0]);
  ai[0] = fmaf(mi, v[0], ai[0]);
  ar[1] = fmaf(mr, v[2], ar[1]);
  ar[1] = fmaf(-mi, v[3], ar[1]);
  ai[1] = fmaf(mr, v[3], ai[1]);
  ai[1] = fmaf(mi, v[2], ai[1]);
}

// N staged entries from p, unmasked (the tail of a whole-step row).
template <int N, int LPX>
__device__ inline void halo_tail(const unsigned char* __restrict__ scol, const float* __restrict__ sval,
                                 const f4* __restrict__ xs, int cl, int p, float* ar, float* ai) {
  int li[N];
  float mr[N], mi[N];
  f4 v[N];
#pragma unroll
  for (int q = 0; q < N; ++q) {
    int iq = p + q;
    asm volatile("" : "+v"(iq));
    li[q] = scol[iq];
  }
#pragma unroll
  for (int q = 0; q < N; ++q) {
    v[q] = xs[li[q] * LPX + cl];
    mr[q] = sval[2 * (p + q)];
    mi[q] = sval[2 * (p + q) + 1];
  }
#pragma unroll
  for (int q = 0; q < N; ++q) cmac2(v[q], mr[q], mi[q], ar, ai);
}

// Row product of the x-halo kernels: ar + i ai = sum over the staged entries
// [s0, s1) of value * x, x read from the LDS halo at slot * LPX + cl.  Four
// entries per step.  Rows of 15 entries (A) take whole steps and an
// unmasked tail of 1-3 entries: the whole steps read their slots at
// constant offsets from p;
// the empty asm keeps each slot read a byte read (merged into one dword at a
// row start that is not 4-aligned, the read measured 13 % slower on the
// A pass).  Shorter rows of varying length (the FSAI factors, at most 8
// entries) measured faster with every step masked.
template <bool WHOLE_STEPS, int LPX>
__device__ inline void halo_row(const unsigned char* __restrict__ scol, const float* __restrict__ sval,
                                const f4* __restrict__ xs, int cl, int s0, int s1, float* ar, float* ai) {
  int p = s0;
  if (WHOLE_STEPS) {
    for (; p + 4 <= s1; p += 4) {
      int li[4];
      float mr[4], mi[4];
      f4 v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        int iq = p + q;
        asm volatile("" : "+v"(iq));
        li[q] = scol[iq];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        v[q] = xs[li[q] * LPX + cl];
        mr[q] = sval[2 * (p + q)];
        mi[q] = sval[2 * (p + q) + 1];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) cmac2(v[q], mr[q], mi[q], ar, ai);
    }
    // the last 1-3 entries without masks (a 15-entry row ends on 3)
    const int rem = s1 - p;
    if (rem == 3) halo_tail<3, LPX>(scol, sval, xs, cl, p, ar, ai);
    else if (rem == 2) halo_tail<2, LPX>(scol, sval, xs, cl, p, ar, ai);
    else if (rem == 1) halo_tail<1, LPX>(scol, sval, xs, cl, p, ar, ai);
    return;
  }
  for (; p < s1; p += 4) {
    int li[4];
    float mr[4], mi[4];
    f4 v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) li[q] = scol[min(p + q, s1 - 1)];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int pq = min(p + q, s1 - 1);
      v[q] = xs[li[q] * LPX + cl];
      mr[q] = sval[2 * pq];
      mi[q] = sval[2 * pq + 1];
      if (p + q >= s1) { mr[q] = 0.f; mi[q] = 0.f; }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) cmac2(v[q], mr[q], mi[q], ar, ai);
  }
}

struct Halo {
  const int* hptr;             // halo chunk -> first entry of its list
  const int* hcol;             // halo lists (columns, mesh-global numbering)
  const unsigned char* hloc;   // nonzero -> slot in its chunk's list
  const int* gchb;             // group -> first halo chunk of its mesh
  const unsigned char* gok;    // group -> its mesh's chunks all fit the LDS halo
  const int* cptr;             // CSR columns (the gather path of meshes that do not fit)
};

// Chunks of a work item in the x-halo kernels.  Contiguous (ilv = 0): the
// item's row tile, chunk after chunk.  Interleaved (ilv = 1): the XCD's
// items of a group (tiles [x*per, x*per + n) of the group, all on XCD x) share
// that tile range's chunks round-robin, item j taking chunks j, j + n,
// j + 2n, ... of the range.  The blocks resident on an XCD then work on
// neighbouring chunks at the same time, so a halo row fetched by one block is
// in L2 for the others; each item still walks up to tile_rows / 64 chunks,
// which keeps the prefetch pipeline.  Chunk s of the item is mesh chunk
// cb + cs * s; nch chunks.
__device__ inline void chunk_map(const Work& W, const Item& it, int ilv, int& cb, int& cs, int& nch) {
  const int grows = W.grows[it.g], T = W.tile_rows, tch = T / kChunkRows;
  const int total = (grows + kChunkRows - 1) / kChunkRows;
  const int t = (it.row0 - W.grow0[it.g]) / T;
  if (!ilv) {
    cb = t * tch;
    cs = 1;
    nch = min(tch, total - cb);
    return;
  }
  // the work list puts tile t on XCD t / per (per = ceil(nt / 8))
  const int nt = (grows + T - 1) / T, per = (nt + 7) / 8;
  const int x = t / per, j = t - x * per;
  const int nreal = min(per, nt - x * per);
  const int cend = min((x * per + nreal) * tch, total);
  cb = x * per * tch + j;
  cs = nreal;
  nch = cb < cend ? (cend - cb + cs - 1) / cs : 0;
}

template <int NSP, int MAT, int EPI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_spmv_halo(
    const int* __restrict__ rptr, const float* __restrict__ v0, const float* __restrict__ v1,
    const float* __restrict__ x, float* __restrict__ y, const float* __restrict__ w,
    const double* __restrict__ coef, const double* __restrict__ coef2, float* __restrict__ xacc,
    double* __restrict__ part1, double* __restrict__ part2, Work W, Halo H, const int* __restrict__ grp_active,
    int ilv) {
  using G = Geo<float, NSP>;
  constexpr int CPL = G::CPL, LPR = G::LPR;
  constexpr int PPT = (kHaloCap * LPR + 255) / 256;   // 16-byte halo pieces per thread
  constexpr int EPT = kChunkCap / 256;                // chunk nonzeros per thread
  constexpr int MAXCH = kHaloTileRows / kChunkRows;
  Item it;
  if (!work_item(W, it)) return;
  if (grp_active && !grp_active[it.gl]) return;
  constexpr int SRS = kChunkRows + 1;   // row pointers per chunk in srp
  __shared__ double sh[2 * 4 * NSP];
  __shared__ int srp[MAXCH * SRS];
  __shared__ int shp[2 * MAXCH];
  __shared__ unsigned char scol[kChunkCap];
  __shared__ float sval[2 * kChunkCap];
  __shared__ f4 xs[kHaloCap * LPR];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int rsub = lane / LPR, cl = lane % LPR, c0 = cl * CPL;
  const int mr0 = W.grow0[it.g], grows = W.grows[it.g];
  int cb, cs, nch;
  chunk_map(W, it, ilv, cb, cs, nch);
  // first row (mesh-relative) and row count of the item's chunk c
  auto crow = [&](int c) __attribute__((always_inline)) { return (cb + cs * c) * kChunkRows; };
  auto cnr = [&](int c) __attribute__((always_inline)) { return min(kChunkRows, grows - crow(c)); };
  float fk = 0.f;
  const float* gv = v0;
  if (MAT == MAT_A) fk = (float)W.gfsc[it.g];
  else gv = v0 + 2 * W.ggv[it.g];
  SpmvEpi<float, NSP, EPI> epi(y, w, xacc, coef, coef2, it.vb, c0, (long long)it.gl * NSP + c0);
  if (!H.gok[it.g]) {
    // a mesh with a chunk whose halo exceeds the LDS: the gather row loop,
    // so that a mesh takes the same path alone and in a batch
    for (int c = 0; c < nch; ++c)
      spmv_rows<float, NSP, MAT, 4>(rptr, H.cptr, gv, v1, fk, x + c0 * 2, it.vb, mr0 + crow(c),
                                    mr0 + crow(c) + cnr(c), srp, reinterpret_cast<int*>(xs), sval, epi);
    epi.finish(sh, part1, part2, it.wl);
    return;
  }
  const f4* xrow = reinterpret_cast<const f4*>(x);   // node n, piece l at n * LPR + l
  for (int q = tid; q < nch * SRS; q += 256) {
    const int c = q / SRS, i = q - c * SRS;
    srp[q] = rptr[mr0 + min(crow(c) + i, grows)];
  }
  for (int c = tid; c < nch; c += 256) {
    const int gc = H.gchb[it.g] + cb + cs * c;
    shp[2 * c] = H.hptr[gc];
    shp[2 * c + 1] = H.hptr[gc + 1];
  }
  __syncthreads();

  // staged state of the next chunk (registers; passed to the stages by
  // reference, which keeps them out of the lambdas' captures and in VGPRs)
  unsigned char ln[EPT];
  float mn0[EPT], mn1[EPT];
  f4 xn[PPT];
  int hcA[PPT], hcB[PPT];
  using XN = f4[PPT];
  using LN = unsigned char[EPT];
  using MN = float[EPT];
  auto load_hcol = [&](int c, int* hc) __attribute__((always_inline)) {
    const int hs = shp[2 * c], npc = (shp[2 * c + 1] - hs) * LPR;
#pragma unroll
    for (int i = 0; i < PPT; ++i) hc[i] = H.hcol[hs + min(tid + 256 * i, npc - 1) / LPR];
  };
  auto load_next = [&](int c, const int* hc, XN& xv, LN& lv, MN& m0v, MN& m1v) __attribute__((always_inline)) {
    const int es = srp[c * SRS], ne = srp[c * SRS + cnr(c)] - es;
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const int q = min(tid + 256 * i, ne - 1);
      lv[i] = H.hloc[es + q];
      if (MAT == MAT_A) {
        m0v[i] = v0[es + q];
        m1v[i] = v1[es + q];
      } else {
        m0v[i] = gv[2 * (long long)(es + q)];
        m1v[i] = gv[2 * (long long)(es + q) + 1];
      }
    }
    const int npc = (shp[2 * c + 1] - shp[2 * c]) * LPR;
#pragma unroll
    for (int i = 0; i < PPT; ++i) {
      const int q = min(tid + 256 * i, npc - 1);
      xv[i] = xrow[(it.vb + hc[i]) * LPR + q % LPR];
    }
  };
  auto store_next = [&](int c, const XN& xv, const LN& lv, const MN& m0v, const MN& m1v)
                        __attribute__((always_inline)) {
    const int es = srp[c * SRS], ne = srp[c * SRS + cnr(c)] - es;
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const int q = min(tid + 256 * i, ne - 1);
      scol[q] = lv[i];
      sval[2 * q] = m0v[i];
      sval[2 * q + 1] = MAT == MAT_A ? fk * m1v[i] : m1v[i];
    }
    const int npc = (shp[2 * c + 1] - shp[2 * c]) * LPR;
#pragma unroll
    for (int i = 0; i < PPT; ++i) xs[min(tid + 256 * i, npc - 1)] = xv[i];
  };
  // epilogue operands of this thread's rows of the current chunk, loaded
  // before the next chunk's prefetch so that waiting for them leaves the
  // prefetch in flight (the vector-memory counter retires in order)
  constexpr int NST = kChunkRows / G::RPB;
  f4 eo1[NST], eo2[NST];
  auto load_ops = [&](int c) __attribute__((always_inline)) {
    const int rb = mr0 + crow(c), nr = cnr(c);
#pragma unroll
    for (int st = 0; st < NST; ++st) {
      const int lr = min(st * G::RPB + wv * G::RPW + rsub, nr - 1);
      epi.load_ops(rb + lr, eo1[st], eo2[st]);
    }
  };
  auto compute = [&](int c) __attribute__((always_inline)) {
    const int rb = mr0 + crow(c), nr = cnr(c);
    const int* cp = srp + c * SRS;
    const int es = cp[0];
#pragma unroll
    for (int st = 0; st < NST; ++st) {
      const int lr = st * G::RPB + wv * G::RPW + rsub;
      if (lr >= nr) continue;
      float ar[CPL], ai[CPL];
#pragma unroll
      for (int e = 0; e < CPL; ++e) { ar[e] = 0.f; ai[e] = 0.f; }
      halo_row<MAT == MAT_A, LPR>(scol, sval, xs, cl, cp[lr] - es, cp[lr + 1] - es, ar, ai);
      epi.apply(rb + lr, ar, ai, eo1[st], eo2[st]);
    }
  };
  // one pipeline step: LDS holds chunk c; hc1 = halo list of c+1, hc2
  // receives that of c+2
  auto step = [&](int c, const int* hc1, int* hc2, XN& xv, LN& lv, MN& m0v, MN& m1v)
                  __attribute__((always_inline)) {
    const bool has1 = c + 1 < nch;
    load_ops(c);
    if (has1) load_next(c + 1, hc1, xv, lv, m0v, m1v);
    if (c + 2 < nch) load_hcol(c + 2, hc2);
    compute(c);
    __syncthreads();
    if (has1) store_next(c + 1, xv, lv, m0v, m1v);
    __syncthreads();
  };
  load_hcol(0, hcA);
  load_next(0, hcA, xn, ln, mn0, mn1);
  if (nch > 1) load_hcol(1, hcB);
  store_next(0, xn, ln, mn0, mn1);
  __syncthreads();
  for (int c = 0; c < nch; c += 2) {
    step(c, hcB, hcA, xn, ln, mn0, mn1);
    if (c + 1 < nch) step(c + 1, hcA, hcB, xn, ln, mn0, mn1);
  }
  epi.finish(sh, part1, part2, it.wl);
}


// Two-half x-halo form for 32-column groups (complex64).  A block walks its
// tile chunk by chunk and each chunk in two halves of 16 columns: the
// chunk's matrix slice is staged in LDS once and serves both halves, while
// the x halo in LDS holds one half at a time (kHaloCap rows x 128 B), so
// three blocks share a CU as with 16-wide groups.  Against 16-wide groups
// the shared A values and slots and the per-frequency G values are read
// once per 32 columns instead of once per 16.  The pipeline runs over
// (chunk, half) sub-steps: while (c, 0) is computed, the half-1 rows of
// chunk c and the halo list of chunk c+1 are in flight; while (c, 1) is
// computed, chunk c+1's matrix slice and half-0 rows are.
template <int MAT, int EPI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_spmv_halo2(
    const int* __restrict__ rptr, const float* __restrict__ v0, const float* __restrict__ v1,
    const float* __restrict__ x, float* __restrict__ y, const float* __restrict__ w,
    const double* __restrict__ coef, const double* __restrict__ coef2, float* __restrict__ xacc,
    double* __restrict__ part1, double* __restrict__ part2, Work W, Halo H, const int* __restrict__ grp_active,
    int ilv) {
  constexpr int NSP = 32, NH = 16, CPL = 2;
  constexpr int LPR = NSP / CPL;        // 16-byte pieces per node row
  constexpr int LPRH = NH / CPL;        // pieces (and lanes) per half row
  constexpr int RPW = 64 / LPRH, RPB = 4 * RPW;
  constexpr int NST = kChunkRows / RPB;
  constexpr int PPT = (kHaloCap * LPRH + 255) / 256;
  constexpr int EPT = kChunkCap / 256;
  constexpr int MAXCH = kHaloTileRows / kChunkRows;
  Item it;
  if (!work_item(W, it)) return;
  if (grp_active && !grp_active[it.gl]) return;
  constexpr int SRS = kChunkRows + 1;   // row pointers per chunk in srp
  __shared__ double sh[2 * 4 * NSP];
  __shared__ int srp[MAXCH * SRS];
  __shared__ int shp[2 * MAXCH];
  __shared__ unsigned char scol[kChunkCap];
  __shared__ float sval[2 * kChunkCap];
  __shared__ f4 xs[kHaloCap * LPRH];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int rsub = lane / LPRH, cl = lane % LPRH;
  const int mr0 = W.grow0[it.g], grows = W.grows[it.g];
  int cb, cs, nch;
  chunk_map(W, it, ilv, cb, cs, nch);
  // first row (mesh-relative) and row count of the item's chunk c
  auto crow = [&](int c) __attribute__((always_inline)) { return (cb + cs * c) * kChunkRows; };
  auto cnr = [&](int c) __attribute__((always_inline)) { return min(kChunkRows, grows - crow(c)); };
  float fk = 0.f;
  const float* gv = v0;
  if (MAT == MAT_A) fk = (float)W.gfsc[it.g];
  else gv = v0 + 2 * W.ggv[it.g];
  SpmvEpi<float, NSP, EPI> ep0(y, w, xacc, coef, coef2, it.vb, cl * CPL, (long long)it.gl * NSP + cl * CPL);
  SpmvEpi<float, NSP, EPI> ep1(y, w, xacc, coef, coef2, it.vb, NH + cl * CPL,
                               (long long)it.gl * NSP + NH + cl * CPL);
  auto finish = [&]() __attribute__((always_inline)) {
    // both halves' column sums into the group's partial slots
    constexpr bool AX = EPI == EPI_AXPBY || EPI == EPI_AXPBY_X;
    if (EPI == EPI_STORE || AX) return;
    auto reduce = [&](Acc<CPL>& a0, Acc<CPL>& a1, double* part) __attribute__((always_inline)) {
#pragma unroll
      for (int q = 0; q < CPL; ++q)
#pragma unroll
        for (int o = LPRH; o < 64; o <<= 1) {
          a0.re[q] += __shfl_xor(a0.re[q], o);
          a0.im[q] += __shfl_xor(a0.im[q], o);
          a1.re[q] += __shfl_xor(a1.re[q], o);
          a1.im[q] += __shfl_xor(a1.im[q], o);
        }
      if (rsub == 0) {
#pragma unroll
        for (int q = 0; q < CPL; ++q) {
          sh[2 * (wv * NSP + cl * CPL + q)] = a0.re[q];
          sh[2 * (wv * NSP + cl * CPL + q) + 1] = a0.im[q];
          sh[2 * (wv * NSP + NH + cl * CPL + q)] = a1.re[q];
          sh[2 * (wv * NSP + NH + cl * CPL + q) + 1] = a1.im[q];
        }
      }
      __syncthreads();
      if (tid < NSP) {
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int ww = 0; ww < 4; ++ww) {
          sr += sh[2 * (ww * NSP + tid)];
          si += sh[2 * (ww * NSP + tid) + 1];
        }
        part[2 * ((long long)it.wl * NSP + tid)] = sr;
        part[2 * ((long long)it.wl * NSP + tid) + 1] = si;
      }
    };
    reduce(ep0.d1, ep1.d1, part1);
    if (EPI == EPI_DOT2 || EPI == EPI_AXPY_SELF || EPI == EPI_STORE_SELF2) {
      __syncthreads();
      reduce(ep0.d2, ep1.d2, part2);
    }
  };
  if (!H.gok[it.g]) {
    // a mesh with a chunk whose halo exceeds the LDS: the gather row loop
    SpmvEpi<float, NSP, EPI> ep(y, w, xacc, coef, coef2, it.vb, (lane % LPR) * CPL,
                                (long long)it.gl * NSP + (lane % LPR) * CPL);
    for (int c = 0; c < nch; ++c)
      spmv_rows<float, NSP, MAT, 4>(rptr, H.cptr, gv, v1, fk, x + (lane % LPR) * CPL * 2, it.vb, mr0 + crow(c),
                                    mr0 + crow(c) + cnr(c), srp, reinterpret_cast<int*>(xs), sval, ep);
    ep.finish(sh, part1, part2, it.wl);
    return;
  }
  const f4* xrow = reinterpret_cast<const f4*>(x);   // node n, piece l at n * LPR + l
  for (int q = tid; q < nch * SRS; q += 256) {
    const int c = q / SRS, i = q - c * SRS;
    srp[q] = rptr[mr0 + min(crow(c) + i, grows)];
  }
  for (int c = tid; c < nch; c += 256) {
    const int gc = H.gchb[it.g] + cb + cs * c;
    shp[2 * c] = H.hptr[gc];
    shp[2 * c + 1] = H.hptr[gc + 1];
  }
  __syncthreads();

  unsigned char ln[EPT];
  float mn0[EPT], mn1[EPT];
  f4 xn[PPT];
  int hcA[PPT], hcB[PPT];
  f4 eo1[NST], eo2[NST];
  using XN = f4[PPT];
  using HC = int[PPT];
  auto load_hcol = [&](int c, HC& hc) __attribute__((always_inline)) {
    const int hs = shp[2 * c], npc = (shp[2 * c + 1] - hs) * LPRH;
#pragma unroll
    for (int i = 0; i < PPT; ++i) hc[i] = H.hcol[hs + min(tid + 256 * i, npc - 1) / LPRH];
  };
  auto load_mat = [&](int c) __attribute__((always_inline)) {
    const int es = srp[c * SRS], ne = srp[c * SRS + cnr(c)] - es;
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const int q = min(tid + 256 * i, ne - 1);
      ln[i] = H.hloc[es + q];
      if (MAT == MAT_A) {
        mn0[i] = v0[es + q];
        mn1[i] = v1[es + q];
      } else {
        mn0[i] = gv[2 * (long long)(es + q)];
        mn1[i] = gv[2 * (long long)(es + q) + 1];
      }
    }
  };
  auto store_mat = [&](int c) __attribute__((always_inline)) {
    const int es = srp[c * SRS], ne = srp[c * SRS + cnr(c)] - es;
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const int q = min(tid + 256 * i, ne - 1);
      scol[q] = ln[i];
      sval[2 * q] = mn0[i];
      sval[2 * q + 1] = MAT == MAT_A ? fk * mn1[i] : mn1[i];
    }
  };
  auto load_x = [&](int c, int h, const HC& hc, XN& xv) __attribute__((always_inline)) {
    const int npc = (shp[2 * c + 1] - shp[2 * c]) * LPRH;
#pragma unroll
    for (int i = 0; i < PPT; ++i) {
      const int q = min(tid + 256 * i, npc - 1);
      xv[i] = xrow[(it.vb + hc[i]) * LPR + h * LPRH + q % LPRH];
    }
  };
  auto store_x = [&](int c, const XN& xv) __attribute__((always_inline)) {
    const int npc = (shp[2 * c + 1] - shp[2 * c]) * LPRH;
#pragma unroll
    for (int i = 0; i < PPT; ++i) xs[min(tid + 256 * i, npc - 1)] = xv[i];
  };
  auto load_ops = [&](int c, SpmvEpi<float, NSP, EPI>& ep) __attribute__((always_inline)) {
    const int rb = mr0 + crow(c), nr = cnr(c);
#pragma unroll
    for (int st = 0; st < NST; ++st) {
      const int lr = min(st * RPB + wv * RPW + rsub, nr - 1);
      ep.load_ops(rb + lr, eo1[st], eo2[st]);
    }
  };
  auto compute = [&](int c, SpmvEpi<float, NSP, EPI>& ep) __attribute__((always_inline)) {
    const int rb = mr0 + crow(c), nr = cnr(c);
    const int* cp = srp + c * SRS;
    const int es = cp[0];
#pragma unroll
    for (int st = 0; st < NST; ++st) {
      const int lr = st * RPB + wv * RPW + rsub;
      if (lr >= nr) continue;
      float ar[CPL], ai[CPL];
#pragma unroll
      for (int e = 0; e < CPL; ++e) { ar[e] = 0.f; ai[e] = 0.f; }
      halo_row<MAT == MAT_A, LPRH>(scol, sval, xs, cl, cp[lr] - es, cp[lr + 1] - es, ar, ai);
      ep.apply(rb + lr, ar, ai, eo1[st], eo2[st]);
    }
  };
  // sub-step (c, 0): LDS holds chunk c's matrix and half-0 rows; hc = c's
  // halo list, hn receives c+1's
  auto half0 = [&](int c, const HC& hc, HC& hn) __attribute__((always_inline)) {
    load_ops(c, ep0);
    load_x(c, 1, hc, xn);
    if (c + 1 < nch) load_hcol(c + 1, hn);
    compute(c, ep0);
    __syncthreads();
    store_x(c, xn);
    __syncthreads();
  };
  // sub-step (c, 1): LDS holds chunk c's matrix and half-1 rows; hn = c+1's
  // halo list
  auto half1 = [&](int c, const HC& hn) __attribute__((always_inline)) {
    const bool has1 = c + 1 < nch;
    load_ops(c, ep1);
    if (has1) {
      load_mat(c + 1);
      load_x(c + 1, 0, hn, xn);
    }
    compute(c, ep1);
    __syncthreads();
    if (has1) {
      store_mat(c + 1);
      store_x(c + 1, xn);
    }
    __syncthreads();
  };
  load_hcol(0, hcA);
  load_mat(0);
  load_x(0, 0, hcA, xn);
  store_mat(0);
  store_x(0, xn);
  __syncthreads();
  for (int c = 0; c < nch; c += 2) {
    half0(c, hcA, hcB);
    half1(c, hcB);
    if (c + 1 < nch) {
      half0(c + 1, hcB, hcA);
      half1(c + 1, hcA);
    }
  }
  finish();
}

// Row loop of the vector passes: every row of the block's work item, LPR
// lanes per row, CPL columns per lane; `off` indexes the row's first column
// of this lane and rhs0 its first RHS in the chunk's per-RHS arrays.
#define ROW_LOOP_BEGIN                                                         \
  using G = Geo<T, NSP>;                                                       \
  constexpr int CPL = G::CPL;                                                  \
  Item it;                                                                     \
  if (!work_item(W, it)) return;                                               \
  if (grp_active && !grp_active[it.gl]) return;                                \
  __shared__ double sh[2 * 4 * NSP];                                           \
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;                    \
  const int rsub = lane / G::LPR, c0 = (lane % G::LPR) * CPL;                  \
  const long long rhs0 = (long long)it.gl * NSP + c0;                          \
  (void)sh; (void)rhs0;

#define ROW_LOOP_ROWS                                                          \
  for (int row = it.row0 + wv * G::RPW + rsub; row < it.row1; row += G::RPB) { \
    const long long off = ((it.vb + row) * NSP + c0) * 2;

#define ROW_LOOP_END }

// s = r - alpha v ; part1 = sum |s|^2
template <typename T, int NSP>
__global__ __launch_bounds__(256) void k_axpy_s(const T* __restrict__ r, const T* __restrict__ v,
                                                T* __restrict__ s, const double* __restrict__ alpha,
                                                double* __restrict__ part1, Work W,
                                                const int* __restrict__ grp_active) {
  ROW_LOOP_BEGIN
  Acc<CPL> d1;
  d1.zero();
  T alr[CPL], ali[CPL];
  load_coef<T, CPL>(alpha, rhs0, alr, ali);
  ROW_LOOP_ROWS
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
  block_reduce_store<T, NSP>(d1, sh, part1, it.wl);
}

// x += alpha h + omega z (active columns only); r = s - omega t;
// part1 = sum rhat0 r, part2 = sum |r|^2
template <typename T, int NSP>
__global__ __launch_bounds__(256) void k_update_xr(
    T* __restrict__ x, const T* __restrict__ h, const T* __restrict__ z, T* __restrict__ r,
    const T* __restrict__ s, const T* __restrict__ t, const T* __restrict__ rhat,
    const double* __restrict__ alpha, const double* __restrict__ omega,
    const int* __restrict__ status, double* __restrict__ part1, double* __restrict__ part2, Work W,
    const int* __restrict__ grp_active) {
  ROW_LOOP_BEGIN
  Acc<CPL> d1, d2;
  d1.zero();
  d2.zero();
  T alr[CPL], ali[CPL], omr[CPL], omi[CPL];
  bool act[CPL];
  load_coef<T, CPL>(alpha, rhs0, alr, ali);
  load_coef<T, CPL>(omega, rhs0, omr, omi);
#pragma unroll
  for (int e = 0; e < CPL; ++e) act[e] = status[rhs0 + e] == ST_ACTIVE;
  ROW_LOOP_ROWS
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
  block_reduce_store<T, NSP>(d1, sh, part1, it.wl);
  __syncthreads();
  block_reduce_store<T, NSP>(d2, sh, part2, it.wl);
}

// p = r + beta (p - omega v)
template <typename T, int NSP>
__global__ __launch_bounds__(256) void k_update_p(T* __restrict__ p, const T* __restrict__ r,
                                                  const T* __restrict__ v, const double* __restrict__ beta,
                                                  const double* __restrict__ omega, Work W,
                                                  const int* __restrict__ grp_active) {
  ROW_LOOP_BEGIN
  T br[CPL], bi[CPL], omr[CPL], omi[CPL];
  load_coef<T, CPL>(beta, rhs0, br, bi);
  load_coef<T, CPL>(omega, rhs0, omr, omi);
  ROW_LOOP_ROWS
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
}

// COCG residual step: r -= alpha q; part2 = sum |r|^2.  (x += alpha p is
// applied in the G^T pass of the same iteration, which reads p anyway.)
template <typename T, int NSP>
__global__ __launch_bounds__(256) void k_cg_update(T* __restrict__ r, const T* __restrict__ q,
                                                   const double* __restrict__ alpha, double* __restrict__ part2,
                                                   Work W, const int* __restrict__ grp_active) {
  ROW_LOOP_BEGIN
  Acc<CPL> d2;
  d2.zero();
  T alr[CPL], ali[CPL];
  load_coef<T, CPL>(alpha, rhs0, alr, ali);
  ROW_LOOP_ROWS
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
  block_reduce_store<T, NSP>(d2, sh, part2, it.wl);
}

// r = b, rhat = b, p = b, x = 0; part1 = sum |b|^2, part2 = sum b b
template <typename T, int NSP>
__global__ __launch_bounds__(256) void k_init(const T* __restrict__ b, T* __restrict__ x,
                                              T* __restrict__ r, T* __restrict__ rhat,
                                              T* __restrict__ p, double* __restrict__ part1,
                                              double* __restrict__ part2, Work W,
                                              const int* __restrict__ grp_active) {
  ROW_LOOP_BEGIN
  Acc<CPL> d1, d2;
  d1.zero();
  d2.zero();
  ROW_LOOP_ROWS
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
  block_reduce_store<T, NSP>(d1, sh, part1, it.wl);
  __syncthreads();
  block_reduce_store<T, NSP>(d2, sh, part2, it.wl);
}

// FP64 refinement residual: r64 = b - A x64 (complex128), with A x64 from
// the staged row products of k_spmv (x64 == nullptr: x64 = 0, r64 = b).
// b is read in the caller's node order when `order` is given (solver row i
// is caller row order[i]), so the right-hand side needs no permuted copy.
// When `write`, stores complex64 r (and p and rhat, if given) = r64 and
// x32 = 0 for columns whose outer status is active (zeros elsewhere);
// part1 = sum |r64|^2 (FP64), part2 = sum r32 r32, part3 (if given) =
// sum |b|^2 (the reference norm when x64 starts from a guess).
template <int NSP>
__global__ __launch_bounds__(256) void k_residual64(
    const int* __restrict__ rptr, const int* __restrict__ cptr, const double* __restrict__ are,
    const double* __restrict__ aim, const double* __restrict__ b, const int* __restrict__ order,
    const double* __restrict__ x64, float* __restrict__ r32, float* __restrict__ rhat32,
    float* __restrict__ p32, float* __restrict__ x32, const int* __restrict__ ostatus, int write,
    double* __restrict__ part1, double* __restrict__ part2, Work W, int ilv, double* __restrict__ part3 = nullptr) {
  using G = Geo<double, NSP>;  // one complex128 per lane
  Item it;
  if (!work_item(W, it)) return;
  __shared__ double sh[2 * 4 * NSP];
  __shared__ int srp[kChunkRows + 1];
  __shared__ int scol[kChunkCap];
  __shared__ double sval[2 * kChunkCap];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int c = lane % G::LPR;
  const long long rhs = (long long)it.gl * NSP + c;
  const bool act = ostatus ? (ostatus[rhs] == ST_ACTIVE) : true;
  // a group none of whose columns is still iterating takes no part: its
  // partial sums are not read and the inner solve skips it
  if (ostatus && !__syncthreads_or(act)) return;
  const long long vb = it.vb;
  Acc<1> d1, d2, d3;
  d1.zero();
  d2.zero();
  d3.zero();
  auto epi = [&](int row, const double* yr, const double* yi) {
    const long long off = (vb + row) * NSP + c;
    const long long boff = order ? (vb + order[row]) * NSP + c : off;
    const double2 bv = *reinterpret_cast<const double2*>(b + 2 * boff);
    d3.re[0] += bv.x * bv.x + bv.y * bv.y;
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
      if (p32) reinterpret_cast<float2*>(p32)[off] = v;
      reinterpret_cast<float2*>(x32)[off] = make_float2(0.f, 0.f);
    }
  };
  if (x64) {
    // the item's chunks as in the x-halo kernels (interleaved: neighbouring
    // chunks on one XCD at a time share their x rows in L2)
    int cb, cs, nch;
    chunk_map(W, it, ilv, cb, cs, nch);
    const int mr0 = W.grow0[it.g], mr1 = mr0 + W.grows[it.g];
    for (int k = 0; k < nch; ++k) {
      const int r0 = mr0 + (cb + cs * k) * kChunkRows;
      spmv_rows<double, NSP, MAT_A, 8>(rptr, cptr, are, aim, W.gfsc[it.g], x64 + c * 2, vb, r0,
                                       min(r0 + kChunkRows, mr1), srp, scol, sval, epi);
    }
  } else {
    const double z[1] = {0.0};
    for (int row = it.row0 + wv * G::RPW + lane / G::LPR; row < it.row1; row += G::RPB) epi(row, z, z);
  }
  block_reduce_store<double, NSP>(d1, sh, part1, it.wl);
  if (write) {
    __syncthreads();
    block_reduce_store<double, NSP>(d2, sh, part2, it.wl);
  }
  if (part3) {
    __syncthreads();
    block_reduce_store<double, NSP>(d3, sh, part3, it.wl);
  }
}

// Initial FP64 iterate of every group of the work list from its guess
// terms: x64[g] = sum_q w[g][q] x64[src[g][q]] (w complex; src -1 ends the terms, no
// term gives 0).  A source is another group of the same mesh and chunk
// (solved in an earlier phase, so not written here), the group itself
// (the caller's x on entry, read before this element is written), or a group
// of an earlier chunk (src < W.g0: its solution is already in the caller's
// x, xout, in the caller's node order).
template <int NSP>
__global__ __launch_bounds__(256) void k_guess(double* __restrict__ x64, Work W, const int* __restrict__ gsrc,
                                               const double* __restrict__ gw, int nterm,
                                               const double* __restrict__ xout, const int* __restrict__ order) {
  Item it;
  if (!work_item(W, it)) return;
  const int nr = it.row1 - it.row0;
  for (int i = threadIdx.x; i < nr * NSP; i += 256) {
    const int row = it.row0 + i / NSP, c = i % NSP;
    double re = 0.0, im = 0.0;
    for (int q = 0; q < nterm; ++q) {
      const int sg = gsrc[(long long)it.g * nterm + q];
      if (sg < 0) break;
      const double2 w = reinterpret_cast<const double2*>(gw)[(long long)it.g * nterm + q];
      double2 v;
      if (sg < W.g0)
        v = reinterpret_cast<const double2*>(xout)[(W.gvb[sg] - W.grow0[sg] + order[row]) * NSP + c];
      else
        v = reinterpret_cast<const double2*>(x64)[(W.gvb[sg] - W.vb0 - W.grow0[sg] + row) * NSP + c];
      re += w.x * v.x - w.y * v.y;
      im += w.x * v.y + w.y * v.x;
    }
    reinterpret_cast<double2*>(x64)[(it.vb + row) * NSP + c] = make_double2(re, im);
  }
}

// x64 = 0 in the columns whose outer status is DOGPUP_RHS_ZERO (a zero
// right-hand side, or a padding column), after a guess wrote them.
template <int NSP>
__global__ __launch_bounds__(256) void k_zero_cols(double* __restrict__ x64, const int* __restrict__ ostatus, Work W) {
  Item it;
  if (!work_item(W, it)) return;
  const int nr = it.row1 - it.row0;
  for (int i = threadIdx.x; i < nr * NSP; i += 256) {
    const int c = i % NSP;
    if (ostatus[(long long)it.gl * NSP + c] != DOGPUP_RHS_ZERO) continue;
    reinterpret_cast<double2*>(x64)[(it.vb + it.row0 + i / NSP) * NSP + c] = make_double2(0.0, 0.0);
  }
}

// x64 += x32 for columns that took part in the sweep.  On the first sweep
// (`first`) x64 holds no values yet: it is written as x32 for those columns
// and 0 for the others, without being read.
template <int NSP>
__global__ __launch_bounds__(256) void k_add_correction(double* __restrict__ x64, const float* __restrict__ x32,
                                                        const int* __restrict__ istatus, int first, Work W) {
  Item it;
  if (!work_item(W, it)) return;
  // a group none of whose columns took part in the sweep is left as it is
  if (!first && !__syncthreads_or((int)threadIdx.x < NSP && istatus[(long long)it.gl * NSP + threadIdx.x] != ST_SKIP))
    return;
  const int nr = it.row1 - it.row0;
  for (int i = threadIdx.x; i < nr * NSP; i += 256) {
    const int c = i % NSP;
    const long long o = (it.vb + it.row0 + i / NSP) * NSP + c;
    const bool skip = istatus[(long long)it.gl * NSP + c] == ST_SKIP;
    if (skip && !first) continue;
    double2 v = make_double2(0.0, 0.0);
    if (!first) v = reinterpret_cast<double2*>(x64)[o];
    if (!skip) {
      const float2 d = reinterpret_cast<const float2*>(x32)[o];
      v.x += d.x;
      v.y += d.y;
    }
    reinterpret_cast<double2*>(x64)[o] = v;
  }
}

// Row permutation of a complex128 block between the caller's node order and
// the solver's: dir 0 gathers (dst[row] = src[order[row]]), dir 1 scatters
// back.  order is global (each mesh's rows map into the same mesh).
template <int NSP>
__global__ __launch_bounds__(256) void k_permute_rows(const double* __restrict__ src, double* __restrict__ dst,
                                                      const int* __restrict__ order, int dir, Work W) {
  Item it;
  if (!work_item(W, it)) return;
  const int nr = it.row1 - it.row0;
  for (int i = threadIdx.x; i < nr * NSP; i += 256) {
    const int row = it.row0 + i / NSP, c = i % NSP;
    const long long a = (it.vb + row) * NSP + c, o = (it.vb + order[row]) * NSP + c;
    const double2 v = reinterpret_cast<const double2*>(src)[dir == 0 ? o : a];
    reinterpret_cast<double2*>(dst)[dir == 0 ? a : o] = v;
  }
}

// ------------------------------------------------ reference-structure passes
// part1 = sum a b (mode 0) or sum |a|^2 (mode 1)
template <typename T, int NSP>
__global__ __launch_bounds__(256) void k_dot(const T* __restrict__ a, const T* __restrict__ bb, int mode,
                                             double* __restrict__ part1, Work W,
                                             const int* __restrict__ grp_active) {
  ROW_LOOP_BEGIN
  Acc<CPL> d1;
  d1.zero();
  ROW_LOOP_ROWS
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
  block_reduce_store<T, NSP>(d1, sh, part1, it.wl);
}

// out = a + op(coef) b; op 0: 1/coef, 1: -1/coef, 2: coef, 3: -coef
template <typename T, int NSP>
__global__ __launch_bounds__(256) void k_axpby(T* __restrict__ out, const T* __restrict__ a,
                                               const T* __restrict__ bb, const double* __restrict__ coef, int op,
                                               Work W, const int* __restrict__ grp_active) {
  ROW_LOOP_BEGIN
  T cr[CPL], ci[CPL];
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
  ROW_LOOP_ROWS
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

// Sum of the per-work-item partials of every column of chunk-local group gl.
// All threads of the (kScalThreads-wide) block call it; thread c < nsp
// receives column c's sum.  Each column is reduced by kScalThreads/nsp
// threads in parallel.
constexpr int kScalThreads = 1024;
__device__ inline void block_sum_part(const double* __restrict__ part, const Work& W, int gl, int nsp,
                                      double& re, double& im) {
  __shared__ double sr[kScalThreads], si[kScalThreads];
  const int t = threadIdx.x, c = t % nsp, q = t / nsp, nq = kScalThreads / nsp;
  const int g = W.g0 + gl, wb = W.gwbeg[g] - W.pw0, wc = W.gwcnt[g];
  // eight loads in flight per thread (one running sum would wait for each)
  constexpr int U = 8;
  double au[U], bu[U];
#pragma unroll
  for (int u = 0; u < U; ++u) { au[u] = 0.0; bu[u] = 0.0; }
  int k = q;
  for (; k + (U - 1) * nq < wc; k += U * nq) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long i = 2 * ((long long)(wb + k + u * nq) * nsp + c);
      au[u] += part[i];
      bu[u] += part[i + 1];
    }
  }
  for (; k < wc; k += nq) {
    const long long i = 2 * ((long long)(wb + k) * nsp + c);
    au[0] += part[i];
    bu[0] += part[i + 1];
  }
  double a = 0.0, b = 0.0;
#pragma unroll
  for (int u = 0; u < U; ++u) { a += au[u]; b += bu[u]; }
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

// Marks chunk-local group gl active if any of its columns is, and counts
// active columns, groups and the groups' rows in s.counter[0..2].
__device__ inline void group_flag(BState s, const Work& W, int gl, bool active) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const unsigned long long m = __ballot(active);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(&cnt, __popcll(m));
  __syncthreads();
  if (threadIdx.x == 0) {
    const int n = cnt;
    s.grp_active[gl] = n > 0;
    if (n) {
      atomicAdd(s.counter, n);
      atomicAdd(s.counter + 1, 1);
      atomicAdd(s.counter + 2, W.grows[W.g0 + gl]);
    }
  }
}

// init from (part1 = |b|^2, part2 = b.b); `mask` (optional outer status)
// leaves non-active columns skipped; columns >= gcols are padding.
__global__ __launch_bounds__(kScalThreads) void k_sc_init(BState s, const double* __restrict__ part1,
                                                          const double* __restrict__ part2,
                                                          const int* __restrict__ mask, int nsp, Work W) {
  const int g = blockIdx.x, c = threadIdx.x;
  double n2 = 0, dummy = 0, rr = 0, ri = 0;
  block_sum_part(part1, W, g, nsp, n2, dummy);
  block_sum_part(part2, W, g, nsp, rr, ri);
  bool active = false;
  if (c < nsp) {
    const long long i = (long long)g * nsp + c;
    const int ncols = W.gcols[W.g0 + g], col = c;
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
  group_flag(s, W, g, active);
}

__global__ __launch_bounds__(kScalThreads) void k_sc_alpha(BState s, const double* __restrict__ part1,
                                                           int nsp, Work W) {
  const int g = blockIdx.x, c = threadIdx.x;
  if (!s.grp_active[g]) return;
  double dr = 0, di = 0;
  block_sum_part(part1, W, g, nsp, dr, di);
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
                                                       Work W) {
  const int g = blockIdx.x, c = threadIdx.x;
  if (!s.grp_active[g]) return;
  double n2 = 0, dummy = 0;
  block_sum_part(part1, W, g, nsp, n2, dummy);
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
                                                           const double* __restrict__ part2, int nsp, Work W) {
  const int g = blockIdx.x, c = threadIdx.x;
  if (!s.grp_active[g]) return;
  double nr = 0, ni = 0, dr = 0, di = 0;
  block_sum_part(part1, W, g, nsp, nr, ni);
  block_sum_part(part2, W, g, nsp, dr, di);
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
                                                          const double* __restrict__ part2, int nsp, Work W,
                                                          int maxit) {
  const int g = blockIdx.x, c = threadIdx.x;
  if (!s.grp_active[g]) return;   // inactive groups stay inactive
  double n2 = 0, dummy = 0, rr = 0, ri = 0;
  block_sum_part(part2, W, g, nsp, n2, dummy);
  block_sum_part(part1, W, g, nsp, rr, ri);
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
  group_flag(s, W, g, active);
}

// ---- COCG scalars (conjugate orthogonal CG: CG with the unconjugated
// bilinear form, for complex-symmetric A and the symmetric M^-1 = G^T G).
// rho = (G r).(G r) = r.M^-1 r, alpha = rho / (p.A p), beta = rho_new / rho.
// A column that is not active gets alpha = beta = 0, so its vectors stay
// bounded while the rest of its group iterates.

// After the start pass u = G r (part1 = u.u): rho = u.u, beta = 0.
// part2 (optional): |u|^2, which then becomes the reference norm of the
// column's relative residual (the u-recurrence form tracks u = G r only).
__global__ __launch_bounds__(kScalThreads) void k_cg_start(BState s, const double* __restrict__ part1,
                                                           const double* __restrict__ part2, int nsp, Work W) {
  const int g = blockIdx.x, c = threadIdx.x;
  double ur = 0, ui = 0, n2 = 0, dummy = 0;
  block_sum_part(part1, W, g, nsp, ur, ui);
  if (part2) block_sum_part(part2, W, g, nsp, n2, dummy);
  bool active = false;
  if (c < nsp) {
    const long long i = (long long)g * nsp + c;
    if (s.status[i] == ST_ACTIVE) {
      if ((ur == 0.0 && ui == 0.0) || !isfinite(ur) || !isfinite(ui) || (part2 && !(n2 > 0.0))) {
        s.status[i] = DOGPUP_RHS_BREAKDOWN;
      } else {
        if (part2) s.bnorm[i] = sqrt(n2);
        s.rho[2 * i] = ur;
        s.rho[2 * i + 1] = ui;
        active = true;
      }
    }
    s.alpha[2 * i] = s.alpha[2 * i + 1] = 0.0;
    s.beta[2 * i] = s.beta[2 * i + 1] = 0.0;
  }
  group_flag(s, W, g, active);
}

// part1 = p.(A p)
__global__ __launch_bounds__(kScalThreads) void k_cg_alpha(BState s, const double* __restrict__ part1,
                                                           int nsp, Work W) {
  const int g = blockIdx.x, c = threadIdx.x;
  if (!s.grp_active[g]) return;
  double dr = 0, di = 0;
  block_sum_part(part1, W, g, nsp, dr, di);
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
                                                          const double* __restrict__ part2, int nsp, Work W,
                                                          int maxit) {
  const int g = blockIdx.x, c = threadIdx.x;
  // grp_prev gates this iteration's G^T pass, which applies x += alpha p
  const int was = s.grp_active[g];
  if (c == 0) s.grp_prev[g] = was;
  if (!was) return;
  double n2 = 0, dummy = 0, ur = 0, ui = 0;
  block_sum_part(part2, W, g, nsp, n2, dummy);
  block_sum_part(part1, W, g, nsp, ur, ui);
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
  group_flag(s, W, g, active);
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
// part3 (sweep 0 only, when x64 starts from a guess): |b|^2, the reference
// norm.  glevel (optional): only groups of level `level` take part (the
// current phase of a continuation solve); the others are left untouched.
__global__ __launch_bounds__(kScalThreads) void k_rf_check(RState o, BState in,
                                                           const double* __restrict__ part1, int nsp, Work W,
                                                           int sweep, double tol, double tol_in, int maxit,
                                                           const double* __restrict__ ctol,
                                                           const double* __restrict__ part3 = nullptr,
                                                           const int* __restrict__ glevel = nullptr,
                                                           int level = 0) {
  const int g = blockIdx.x, c = threadIdx.x;
  if (glevel && glevel[W.g0 + g] != level) {
    group_flag(in, W, g, false);
    return;
  }
  double n2 = 0, dummy = 0, b2 = 0;
  block_sum_part(part1, W, g, nsp, n2, dummy);
  if (part3) block_sum_part(part3, W, g, nsp, b2, dummy);
  bool active = false;
  if (c < nsp) {
    const long long i = (long long)g * nsp + c;
    const int ncols = W.gcols[W.g0 + g], col = c;
    if (sweep == 0) {
      if (!part3) b2 = n2;
      o.bnorm[i] = sqrt(b2);
      o.iters[i] = 0;
      o.prev[i] = INFINITY;
      o.status[i] = (col >= ncols || b2 == 0.0) ? DOGPUP_RHS_ZERO : ST_ACTIVE;
      o.relres[i] = (o.status[i] == ST_ACTIVE) ? 1.0 : 0.0;
    }
    if (o.status[i] == ST_ACTIVE) {
      const double tl = (ctol && ctol[i] > 0.0) ? ctol[i] : tol;
      const double rel = sqrt(n2) / o.bnorm[i];
      o.relres[i] = rel;
      if (rel < tl) o.status[i] = DOGPUP_RHS_CONVERGED;
      else if (!isfinite(rel)) o.status[i] = DOGPUP_RHS_BREAKDOWN;
      else if (o.iters[i] >= maxit) o.status[i] = DOGPUP_RHS_MAXIT;
      else if (rel > 0.5 * o.prev[i]) o.status[i] = DOGPUP_RHS_BREAKDOWN;  // stagnation
      else {
        o.prev[i] = rel;
        double t = 0.5 * tl / rel;
        if (t < tol_in) t = tol_in;
        if (t > 0.5) t = 0.5;
        in.tol[i] = t;
        active = true;
      }
    }
  }
  group_flag(in, W, g, active);
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
                                                             const double* __restrict__ part2, int nsp, Work W,
                                                             unsigned long long* __restrict__ maxbits) {
  const int g = blockIdx.x, c = threadIdx.x;
  double ar = 0, ai = 0, br = 0, bi = 0;
  block_sum_part(part1, W, g, nsp, ar, ai);
  if (part2) block_sum_part(part2, W, g, nsp, br, bi);
  if (c >= nsp) return;
  const long long i = (long long)g * nsp + c;
  const int ncols = W.gcols[W.g0 + g], col = c;
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

// ------------------------------------------ reduced-basis initial guesses --
// Galerkin continuation.  The converged solutions x_j of one (mesh, column
// block) at the frequencies already solved span a space that holds the
// solution at a nearby frequency k to high accuracy (the diffusion fields
// are smooth in the modulation frequency).  The guess of column c at
// frequency k is the Galerkin solution of A_k x = b_k in the span of the
// kRbTerms solutions nearest in frequency: W^H A_k W z = W^H b_k with W an
// M-orthonormal basis of the span.  For point sources b_j = s_j q (the
// source spectrum s_j times a fixed vector q, dMesh.m:128,196) and
// A_j = K + i sigma_j M, so A_j x_j = b_j gives K x_j = s_j q - i sigma_j M x_j
// and every entry of the small system follows from two kinds of dots:
//   x_i^H A_k x_j = s_j (x_i^H q) + i (sigma_k - sigma_j) (x_i^H M x_j).
// The kernels below compute M x_j with x_j^H b_j (k_rb_mx, k_rb_csum), the
// column-wise M-Gram entries (k_rb_gram, k_rb_reduce), the weights of every
// target column (k_rb_weights) and the guesses (k_rb_guess).
constexpr int kRbTerms = 6;            // solutions in a guess's window
constexpr int kRbPairsPerThread = 24;  // Gram pairs accumulated per thread
constexpr int kRbTgtPerThread = 4;     // target groups per thread of k_rb_guess
constexpr int kRbBlocks = 512;         // blocks of the Gram and guess launches

// mx = M x64 for every group of the work list (M = the unit-frequency
// imaginary part of A, real), stored at x64's offsets, and part1 = sum over
// the item's rows of conj(x64) b per column (b in the caller's node order).
template <int NSP>
__global__ __launch_bounds__(256) void k_rb_mx(const int* __restrict__ rptr, const int* __restrict__ cptr,
                                               const double* __restrict__ mval, const double* __restrict__ x64,
                                               double* __restrict__ mx, const double* __restrict__ b,
                                               const int* __restrict__ order, double* __restrict__ part1, Work W,
                                               int ilv) {
  using G = Geo<double, NSP>;
  Item it;
  if (!work_item(W, it)) return;
  __shared__ double sh[2 * 4 * NSP];
  __shared__ int srp[kChunkRows + 1];
  __shared__ int scol[kChunkCap];
  __shared__ double sval[2 * kChunkCap];
  const int c = (threadIdx.x & 63) % G::LPR;
  const long long vb = it.vb;
  Acc<1> d1;
  d1.zero();
  auto epi = [&](int row, const double* yr, const double* yi) {
    const long long off = (vb + row) * NSP + c;
    reinterpret_cast<double2*>(mx)[off] = make_double2(yr[0], yi[0]);
    const double2 xv = reinterpret_cast<const double2*>(x64)[off];
    const double2 bv = reinterpret_cast<const double2*>(b)[(vb + order[row]) * NSP + c];
    d1.re[0] += xv.x * bv.x + xv.y * bv.y;
    d1.im[0] += xv.x * bv.y - xv.y * bv.x;
  };
  int cb, cs, nch;
  chunk_map(W, it, ilv, cb, cs, nch);
  const int mr0 = W.grow0[it.g], mr1 = mr0 + W.grows[it.g];
  for (int k = 0; k < nch; ++k) {
    const int r0 = mr0 + (cb + cs * k) * kChunkRows;
    // MAT_A with scale 0: the values of M alone
    spmv_rows<double, NSP, MAT_A, 8>(rptr, cptr, mval, mval, 0.0, x64 + c * 2, vb, r0, min(r0 + kChunkRows, mr1),
                                     srp, scol, sval, epi);
  }
  block_reduce_store<double, NSP>(d1, sh, part1, it.wl);
}

// out[g*nsp + c] = column c's sum of part1 over the items of chunk-local
// group g = glist[i] (one block per listed group).
__global__ __launch_bounds__(kScalThreads) void k_rb_csum(const double* __restrict__ part1,
                                                          double* __restrict__ out, const int* __restrict__ glist,
                                                          int nsp, Work W) {
  const int g = glist[blockIdx.x];
  double re = 0.0, im = 0.0;
  block_sum_part(part1, W, g, nsp, re, im);
  if ((int)threadIdx.x < nsp) {
    out[2 * ((long long)g * nsp + threadIdx.x)] = re;
    out[2 * ((long long)g * nsp + threadIdx.x) + 1] = im;
  }
}

// Weights of every (target, column) of a phase, one thread each, from the
// Gram entries and the dots x^H b (rb_weights).  Per target t, tinfo holds
// 2 + kRbTerms^2 + kRbTerms ints: the window size n, the columns holding a
// right-hand side, the code 2 * pair + conjugate of Gram entry (a, b) (-1
// outside the window) and each window solution's chunk-local group; tspec
// holds 3 kRbTerms + 3 doubles: the spectra (complex) and scales of the
// window, then the target's spectrum and scale.  w[(t*kRbTerms + i)*nsp + c]
// receives weight i (zero for padding columns).
__global__ __launch_bounds__(64) void k_rb_weights(const int* __restrict__ tinfo, const double* __restrict__ tspec,
                                                   const double* __restrict__ gram, const double* __restrict__ cdot,
                                                   double* __restrict__ w, int ntgt, int nsp) {
  constexpr int K = kRbTerms, TI = 2 + K * K + K, TS = 3 * K + 3;
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (long long)ntgt * nsp) return;
  const int t = (int)(id / nsp), c = (int)(id % nsp);
  const int* ti = tinfo + (size_t)t * TI;
  const double* ts = tspec + (size_t)t * TS;
  const int n = ti[0];
  double G[2 * K * K], cq[2 * K], y[2 * K];
  for (int i = 0; i < 2 * K; ++i) y[i] = 0.0;
  if (c < ti[1] && n > 0) {
    for (int a = 0; a < n; ++a) {
      // x^H q = x^H b / s (the window's spectra are nonzero)
      const double* d = cdot + 2 * ((long long)ti[2 + K * K + a] * nsp + c);
      const RbC q = rb_div(RbC{d[0], d[1]}, RbC{ts[2 * a], ts[2 * a + 1]});
      cq[2 * a] = q.re;
      cq[2 * a + 1] = q.im;
      for (int b = 0; b < n; ++b) {
        const int code = ti[2 + a * K + b];
        const double* e = gram + 2 * ((long long)(code >> 1) * nsp + c);
        G[2 * (a * n + b)] = e[0];
        G[2 * (a * n + b) + 1] = (code & 1) ? -e[1] : e[1];
      }
    }
    rb_weights<K>(n, G, cq, ts, ts + 2 * K, ts[3 * K], ts[3 * K + 1], ts[3 * K + 2], y);
  }
  for (int i = 0; i < K; ++i) {
    w[2 * (((long long)t * K + i) * nsp + c)] = y[2 * i];
    w[2 * (((long long)t * K + i) * nsp + c) + 1] = y[2 * i + 1];
  }
}

// Row staging of the Gram and guess kernels.  A tile is R consecutive rows
// (mesh-relative) of nl listed vectors, [list][row][c] in LDS, at most
// kRbStage double2 per thread: the element e = threadIdx.x + 256 i of list
// e / (R NSP) sits at a fixed offset from the tile's first row, so every
// thread keeps its kRbStage source pointers and issues the next tile's
// loads (all in flight) before the current tile's products.
constexpr int kRbStage = 8;
constexpr int kRbLds = 256 * kRbStage * 16;   // LDS bytes of a tile
template <int NSP>
struct RbStage {
  const double2* src[kRbStage];   // column of the element in row 0 of its vector
  int row[kRbStage];              // row within the tile; -1: no element
  double2 v[kRbStage];
  // vector u of the list starts at base(u) (complex elements, row 0)
  template <typename Base>
  __device__ inline void init(int nl, int R, Base base) {
    const int per = R * NSP;
#pragma unroll
    for (int i = 0; i < kRbStage; ++i) {
      const int e = threadIdx.x + 256 * i;
      row[i] = -1;
      src[i] = nullptr;
      if (e < nl * per) {
        const int u = e / per, k = e - u * per;
        row[i] = k / NSP;
        src[i] = base(u) + (k - row[i] * NSP);
      }
    }
  }
  // the tile of rows [r0, r0 + nr); an element past the mesh's last row
  // reads row r0 instead (in bounds, never used)
  __device__ inline void load(int r0, int nr) {
#pragma unroll
    for (int i = 0; i < kRbStage; ++i) {
      const int r = (row[i] >= 0 && row[i] < nr) ? row[i] : 0;
      v[i] = src[i] ? src[i][((long long)r0 + r) * NSP] : make_double2(0.0, 0.0);
    }
  }
  __device__ inline void store(double2* lds) const {
#pragma unroll
    for (int i = 0; i < kRbStage; ++i)
      if (row[i] >= 0) lds[threadIdx.x + 256 * i] = v[i];
  }
};

// Column-wise Gram entries over the nrows rows of one mesh: pair p gives
// part[(blk*np + p)*NSP + c] = sum over this block's rows of
// conj(x64[offa[pia[p]] + row][c]) * mx[offb[pib[p]] + row][c].  Tiles of R
// rows, tile t = blockIdx.x + k*gridDim.x, staged as above (the na listed
// x64 vectors, then the nb listed mx vectors).
template <int NSP>
__global__ __launch_bounds__(256) void k_rb_gram(const double* __restrict__ x64, const double* __restrict__ mx,
                                                 const long long* __restrict__ offa, int na,
                                                 const long long* __restrict__ offb, int nb,
                                                 const int* __restrict__ pia, const int* __restrict__ pib, int np,
                                                 int nrows, int R, double* __restrict__ part) {
  extern __shared__ double2 rb_lds[];
  constexpr int PQ = 256 / NSP;
  // LDS offsets (in double2, row 0, column 0) of each pair's two vectors
  __shared__ int lpa[PQ * kRbPairsPerThread], lpb[PQ * kRbPairsPerThread];
  const int c = threadIdx.x % NSP, q = threadIdx.x / NSP;
  for (int p = threadIdx.x; p < np; p += 256) {
    lpa[p] = pia[p] * R * NSP;
    lpb[p] = (na + pib[p]) * R * NSP;
  }
  double ar[kRbPairsPerThread], ai[kRbPairsPerThread];
  const int npt = np > q ? (np - q + PQ - 1) / PQ : 0;   // this thread's pairs
#pragma unroll
  for (int i = 0; i < kRbPairsPerThread; ++i) {
    ar[i] = 0.0;
    ai[i] = 0.0;
  }
  RbStage<NSP> stg;
  stg.init(na + nb, R, [&](int u) {
    return reinterpret_cast<const double2*>(u < na ? x64 : mx) + (u < na ? offa[u] : offb[u - na]) * NSP;
  });
  const int ntile = (nrows + R - 1) / R;
  int t = blockIdx.x;
  if (t < ntile) stg.load(t * R, min(R, nrows - t * R));
  for (; t < ntile; t += gridDim.x) {
    const int nr = min(R, nrows - t * R);
    __syncthreads();   // the previous tile's products are done
    stg.store(rb_lds);
    __syncthreads();
    const int tn = t + gridDim.x;
    if (tn < ntile) stg.load(tn * R, min(R, nrows - tn * R));
    for (int r = 0; r < nr; ++r) {
#pragma unroll
      for (int i = 0; i < kRbPairsPerThread; ++i) {
        if (i < npt) {
          const int p = q + i * PQ;
          const double2 a = rb_lds[lpa[p] + r * NSP + c];
          const double2 bb = rb_lds[lpb[p] + r * NSP + c];
          ar[i] += a.x * bb.x + a.y * bb.y;
          ai[i] += a.x * bb.y - a.y * bb.x;
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < kRbPairsPerThread; ++i) {
    if (i < npt) {
      const long long o = ((long long)blockIdx.x * np + q + i * PQ) * NSP + c;
      part[2 * o] = ar[i];
      part[2 * o + 1] = ai[i];
    }
  }
}

// out[i] = sum over nblk blocks of part[blk*n + i] (complex), i < n.
__global__ void k_rb_reduce(const double* __restrict__ part, double* __restrict__ out, int n, int nblk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr int U = 8;   // eight loads in flight
  double sr[U], si[U];
#pragma unroll
  for (int u = 0; u < U; ++u) { sr[u] = 0.0; si[u] = 0.0; }
  int b = 0;
  for (; b + U <= nblk; b += U)
#pragma unroll
    for (int u = 0; u < U; ++u) {
      sr[u] += part[2 * ((long long)(b + u) * n + i)];
      si[u] += part[2 * ((long long)(b + u) * n + i) + 1];
    }
  for (; b < nblk; ++b) {
    sr[0] += part[2 * ((long long)b * n + i)];
    si[0] += part[2 * ((long long)b * n + i) + 1];
  }
  double r = 0.0, m = 0.0;
#pragma unroll
  for (int u = 0; u < U; ++u) { r += sr[u]; m += si[u]; }
  out[2 * i] = r;
  out[2 * i + 1] = m;
}

// Guesses of nt target groups of one mesh: x64[offt[t] + row][c] =
// sum_{i < kRbTerms} w[(t*kRbTerms + i)*NSP + c] * x64[offa[src[t*kRbTerms + i]] + row][c]
// (src -1: no term; na = 0 writes zeros).  The rows of the na source
// solutions (none of them a target) are staged as in k_rb_gram.
template <int NSP>
__global__ __launch_bounds__(256) void k_rb_guess(double* __restrict__ x64, const long long* __restrict__ offa,
                                                  int na, const long long* __restrict__ offt, int nt,
                                                  const int* __restrict__ src, const double* __restrict__ w,
                                                  int nrows, int R) {
  extern __shared__ double2 rb_lds[];
  constexpr int PQ = 256 / NSP;
  const int c = threadIdx.x % NSP, q = threadIdx.x / NSP;
  int si[kRbTgtPerThread][kRbTerms];
  double2 wt[kRbTgtPerThread][kRbTerms];
  long long ot[kRbTgtPerThread];
#pragma unroll
  for (int j = 0; j < kRbTgtPerThread; ++j) {
    const int t = min(q + j * PQ, nt - 1);
    ot[j] = offt[t];
#pragma unroll
    for (int i = 0; i < kRbTerms; ++i) {
      const int s = src[t * kRbTerms + i];
      si[j][i] = max(s, 0);
      wt[j][i] = s >= 0 ? reinterpret_cast<const double2*>(w)[((long long)t * kRbTerms + i) * NSP + c]
                        : make_double2(0.0, 0.0);
    }
  }
  const int ntt = nt > q ? (nt - q + PQ - 1) / PQ : 0;   // this thread's targets
  RbStage<NSP> stg;
  stg.init(na, R, [&](int u) { return reinterpret_cast<const double2*>(x64) + offa[u] * NSP; });
  const int ntile = (nrows + R - 1) / R;
  int t = blockIdx.x;
  if (t < ntile) stg.load(t * R, min(R, nrows - t * R));
  for (; t < ntile; t += gridDim.x) {
    const int r0 = t * R, nr = min(R, nrows - r0);
    __syncthreads();
    stg.store(rb_lds);
    __syncthreads();
    const int tn = t + gridDim.x;
    if (tn < ntile) stg.load(tn * R, min(R, nrows - tn * R));
    for (int r = 0; r < nr; ++r) {
#pragma unroll
      for (int j = 0; j < kRbTgtPerThread; ++j) {
        if (j < ntt) {
          double re = 0.0, im = 0.0;
          if (na > 0) {
#pragma unroll
            for (int i = 0; i < kRbTerms; ++i) {
              const double2 v = rb_lds[(si[j][i] * R + r) * NSP + c];
              re += wt[j][i].x * v.x - wt[j][i].y * v.y;
              im += wt[j][i].x * v.y + wt[j][i].y * v.x;
            }
          }
          reinterpret_cast<double2*>(x64)[(ot[j] + r0 + r) * NSP + c] = make_double2(re, im);
        }
      }
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
  int nmesh = 0, nn = 0, nsp = 0, ngroups = 0, nwork = 0, mode = 0;
  int nnzA = 0, nnzG = 0, nrhs_out = 0, nf_total = 0;
  int method = DOGPUP_METHOD_COCG;  // Krylov method of the fused forms (the plain form: BiCGSTAB)
  // meshes: rows [m_row0[i], m_row0[i+1]), nf, ncols, column blocks, first
  // group, first A and G entries, G entries, G value base (complex elements,
  // [f][nnzG_i] per mesh), first fscale entry, first output RHS
  std::vector<int> m_row0, m_nf, m_ncols, m_ncb, m_g0, m_pA0, m_pG0, m_nnzG, m_fs0, m_rhs0;
  std::vector<long long> m_gbase;
  long long gvals = 0;  // complex G values per copy (sum of nf_i * nnzG_i)
  // groups (host copies of the device table)
  // h_gibeg: first work item of each group (items padded to multiples of
  // 8); h_gwbeg/h_gwcnt: its partial-sum slots (one per row tile)
  std::vector<int> h_grow0, h_grows, h_gcols, h_gmesh, h_gf, h_gcb, h_gibeg, h_gwbeg, h_gwcnt;
  std::vector<int> h_wgrp, h_wtile, h_wpart;   // host copy of the work list
  std::vector<long long> h_gvb, h_ggv;
  int nparts = 0;
  dogpup::Buf t_wgrp, t_wtile, t_wpart, t_grow0, t_grows, t_gvb, t_gfsc, t_ggv, t_gwbeg, t_gwcnt, t_gcols;
  std::vector<int> chunk_g;  // chunk c = groups [chunk_g[c], chunk_g[c+1])
  // caller's patterns in the reference node order (FSAI setup runs on these)
  const int *rA0 = nullptr, *cA0 = nullptr, *rG0 = nullptr, *cG0 = nullptr;
  // patterns in solver node order (owned): A, G, G^T and the value maps
  // slot -> slot of the reference-order arrays
  const int *rA = nullptr, *cA = nullptr, *rG = nullptr, *cG = nullptr;
  const int *rGT = nullptr, *cGT = nullptr;
  dogpup::Buf pAr, pAc, pGr, pGc, pGTr, pGTc, amap, gmap, gtmap, order;
  // the caller's node order is the solver's (a mesh already on its Morton
  // curve): the solves then work in the caller's b and x, with no row
  // permutation in or out, unless b and x overlap (ident_solve: this solve
  // does).  The FP64 fused form's captured iteration then updates the
  // caller's x in place, so its graphs hold that address (graph_x).
  bool order_identity = false, ident_solve = false;
  const double* graph_x = nullptr;
  // x-halo lists of the three solver-order patterns (A, G, G^T) per 64-row
  // chunk of each mesh (k_spmv_halo); halo_ok[m]: every chunk of pattern m
  // fits the LDS halo; use_halo: the complex64 SpMVs take the halo form
  dogpup::Buf h_ptr[3], h_col[3], h_loc[3], t_gok[3], t_gchb;
  bool halo_ok[3] = {false, false, false};   // some mesh of the handle fits
  std::vector<unsigned char> mesh_ok_tmp[3];  // per mesh (while building)
  bool use_halo = true;
  bool halo_halves = true;   // 32-wide groups: the two-half halo kernel
  bool halo_ilv = true;      // halo SpMVs, FP64 residual: chunks interleaved over an XCD's items
  bool cg_rfree = true;   // the refinement form's inner COCG updates u = G r directly
  std::vector<int> mesh_chb_tmp;   // first halo chunk of each mesh (while building)
  // values: a64re/a64im and g64 in reference order; the rest in solver order
  dogpup::Buf a64re, a64im, a64pre, a64pim, a32re, a32im, fscale, g64, gp64, gtp64, g32, gt32;
  // FSAI straight into the solver layouts (every pattern row of at most 8
  // entries): A positions of each row's lower triangle on its pattern (36 per
  // row), each reference-order G entry's position in the G layout (ginv) and
  // each G^T entry's position in the G layout (gtg); g64 (reference order,
  // FP64) is then filled only on request
  bool fsai_direct = false, g64_valid = false;
  dogpup::Buf fsai_pos, ginv, gtg;
  dogpup::Buf ws, part1, part2, part3, state, rstate, bperm, xperm;
  int tile_rows = 512, lds_pad = 0;
  int unroll_a = 8, unroll_g = 4;  // gathers in flight per row (k_spmv: 4, 8 or 16)
  double ws_budget = 0;             // workspace budget, fixed when the handle is made
  // per-RHS tolerances in group-slot order (g * nsp + c); 0 = the solve's tol
  dogpup::Buf rhs_tol;
  bool have_rhs_tol = false;
  // Initial guesses (refinement form).  Per frequency (global over the
  // meshes) up to guess_nterm (source frequency, weight) terms, as given by
  // dogpup_solver_set_guess or built from fscale by the frequency
  // continuation schedule (freq_cont); per group the same terms as source
  // groups (h_gsrc, -1 ends them; the group itself = the caller's x on
  // entry) and its level (phase).  Each chunk runs its groups level by level
  // over a work list permuted so that every level is one contiguous range.
  int guess_nterm = 0;
  std::vector<int> guess_fsrc;       // nf_total * guess_nterm
  std::vector<double> guess_fw;      // complex: 2 * nf_total * guess_nterm
  std::vector<double> h_spec;        // right-hand-side spectrum of the schedule (complex per frequency; empty: flat)
  bool freq_cont = false;            // schedule from fscale (rebuilt at set_matrix)
  bool phases_dirty = true;
  struct Phase {
    int level, w0, nw;
    bool guess;        // some group of the phase starts from a guess
    long long nodes;   // rows summed over the phase's groups
  };
  std::vector<std::vector<Phase>> chunk_phases;
  bool guess_self = false;           // some group reads the caller's x
  std::vector<double> h_fscale;      // host copy of fscale (set_matrix)
  dogpup::Buf t_pwgrp, t_pwtile, t_pwpart, t_gsrc, t_gw, t_glevel;
  // Galerkin guesses of the continuation schedule (cont_rb, the default;
  // else Lagrange): host copies of the guess terms and of the phase work
  // list, and per chunk whether every term lies in the chunk (else the
  // chunk takes the Lagrange guesses)
  bool cont_rb = true;
  int cont_stride = 0;               // level-0 spacing; 0: the default of the guess kind
  std::vector<int> h_gsrc, h_pwgrp;
  std::vector<unsigned char> chunk_rb;
  // the Galerkin plan of every chunk and phase (rb_build_plan): launch
  // descriptors with offsets into rb_plan (the uploaded lists and tables),
  // the per-group dots x^H b (rb_cdot), the Gram entries (rb_gram), the
  // weights of one phase (rb_w) and the Gram kernel's block partials
  struct RbGram {
    size_t o_offa = 0, o_offb = 0, o_pia = 0, o_pib = 0;
    int na = 0, nb = 0, np = 0, nrows = 0, R = 1;
    long long out = 0;     // first pair id of the launch
  };
  struct RbGuess {
    size_t o_offa = 0, o_offt = 0, o_src = 0;
    int na = 0, nt = 0, nrows = 0, R = 1, w0 = 0;   // w0: first target in the phase's weights
  };
  struct RbPhase {
    bool ok = false;
    std::vector<int> mx_phases;   // phases whose solutions get M x and x^H b first
    size_t o_glist = 0, o_tinfo = 0, o_tspec = 0;
    int nglist = 0, ntgt = 0;
    std::vector<RbGram> gram;
    std::vector<RbGuess> guess;
  };
  std::vector<std::vector<RbPhase>> rb_phase;
  dogpup::Buf rb_plan, rb_cdot, rb_gram, rb_w, rb_part;
  bool have_matrix = false;
  double stats[4] = {0, 0, 0, 0};
  int* h_counter = nullptr;  // pinned: [0] active RHS, [1] active groups, [2] their rows
  unsigned long long* h_max = nullptr;
  int active_groups = 0;
  long long active_rows = 0;
  // optional per-launch timing: event pairs resolved after the solve
  struct Rec {
    hipEvent_t a, b;
    int cls;
    double bytes;
  };
  bool profile = false;
  std::vector<Rec> recs;
  // captured Krylov iterations, keyed by (chunk, maxit, method, precision)
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

// out[p] = in[map[p]] (real values, type conversion)
template <typename TO>
__global__ void k_gather_real(const double* __restrict__ in, const int* __restrict__ map,
                              TO* __restrict__ out, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = (TO)in[map[i]];
}

// One mesh's G values into solver order: for p in [p0, p0 + nnz) and every
// frequency f < nf, out[base + f*nnz + p - p0] = in[base + f*nnz + map[p] - p0]
// (complex; map stays within the mesh).
template <typename TI, typename TO>
__global__ void k_gather_cplx(const TI* __restrict__ in, const int* __restrict__ map, TO* __restrict__ out,
                              long long p0, long long nnz, int nf, long long base) {
  const long long total = nnz * nf;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const long long f = i / nnz, p = i % nnz;
    const long long d = base + f * nnz + p;
    const long long s = base + f * nnz + (map[p0 + p] - p0);
    out[2 * d] = (TO)in[2 * s];
    out[2 * d + 1] = (TO)in[2 * s + 1];
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

// The group table and work list of chunk c.
static Work chunk_work(const dogpup_solver* S, int c) {
  Work W;
  W.wgrp = S->t_wgrp.as<int>();
  W.wtile = S->t_wtile.as<int>();
  W.wpart = S->t_wpart.as<int>();
  W.grow0 = S->t_grow0.as<int>();
  W.grows = S->t_grows.as<int>();
  W.gvb = S->t_gvb.as<long long>();
  W.gfsc = S->t_gfsc.as<double>();
  W.ggv = S->t_ggv.as<long long>();
  W.gwbeg = S->t_gwbeg.as<int>();
  W.gwcnt = S->t_gwcnt.as<int>();
  W.gcols = S->t_gcols.as<int>();
  const int g0 = S->chunk_g[c], g1 = S->chunk_g[c + 1];
  W.g0 = g0;
  W.w0 = S->h_gibeg[g0];
  W.nw = (g1 < S->ngroups ? S->h_gibeg[g1] : S->nwork) - W.w0;
  W.pw0 = S->h_gwbeg[g0];
  W.vb0 = S->h_gvb[g0];
  W.tile_rows = S->tile_rows;
  return W;
}

// nodes (rows summed over groups) of groups [g0, g1)
static long long group_nodes(const dogpup_solver* S, int g0, int g1) {
  return (g1 < S->ngroups ? S->h_gvb[g1] : S->h_gvb.back() + S->h_grows.back()) - S->h_gvb[g0];
}

template <typename T, int NSP>
struct Chunk {
  dogpup_solver* S;
  hipStream_t st;
  int cidx, gcount;
  int phase = 0;   // phase of a continuation solve (its own work list and graphs)
  Work W;
  BState bs;
  double *p1, *p2;
  long long nodes;  // nodes of all groups of the chunk

  dim3 vgrid() const { return dim3((unsigned)W.nw); }
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
  int launch(const int* rp, const int* cp, const T* m0, const T* m1, const T* x, T* y, const T* w, const int* ga,
             const double* coef, const double* coef2, T* xacc) {
    DP_TRY((pad_lds<MAT, EPI, U>()));
    k_spmv<T, NSP, MAT, EPI, U><<<vgrid(), 256, S->lds_pad, st>>>(rp, cp, m0, m1, x, y, w, coef, coef2, xacc, p1,
                                                                  p2, W, ga);
    DP_LAUNCH_CHECK();
    return DOGPUP_OK;
  }
  template <int MAT, int EPI>
  int launch_u(int u, const int* rp, const int* cp, const T* m0, const T* m1, const T* x, T* y, const T* w,
               const int* ga, const double* coef, const double* coef2, T* xacc) {
    switch (u) {
      case 4: return launch<MAT, EPI, 4>(rp, cp, m0, m1, x, y, w, ga, coef, coef2, xacc);
      case 8: return launch<MAT, EPI, 8>(rp, cp, m0, m1, x, y, w, ga, coef, coef2, xacc);
      default: return launch<MAT, EPI, 16>(rp, cp, m0, m1, x, y, w, ga, coef, coef2, xacc);
    }
  }
  // the x-halo form (complex64, NSP <= 32) when the pattern's chunks fit it
  static constexpr bool kHaloType = std::is_same<T, float>::value && NSP <= 32;
  // (the halo kernel holds its tile's row pointers in LDS: tiles of at most
  // kHaloTileRows rows)
  bool halo(int m) const {
    return kHaloType && S->use_halo && S->halo_ok[m] && S->tile_rows <= kHaloTileRows;
  }
  template <int MAT, int EPI>
  int launch_halo(int m, const int* rp, const T* m0, const T* m1, const T* x, T* y, const T* w, const int* ga,
                  const double* coef, const double* coef2, T* xacc) {
    if constexpr (kHaloType) {
      Halo H{S->h_ptr[m].template as<int>(), S->h_col[m].template as<int>(),
             S->h_loc[m].template as<unsigned char>(), S->t_gchb.template as<int>(),
             S->t_gok[m].template as<unsigned char>(), m == 0 ? S->cA : (m == 1 ? S->cG : S->cGT)};
      if constexpr (NSP == 32) {
        if (S->halo_halves) {
          if (S->lds_pad > 0)
            DP_HIP(hipFuncSetAttribute((const void*)k_spmv_halo2<MAT, EPI>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, S->lds_pad));
          k_spmv_halo2<MAT, EPI><<<vgrid(), 256, S->lds_pad, st>>>(rp, m0, m1, x, y, w, coef, coef2, xacc, p1, p2, W,
                                                                    H, ga, S->halo_ilv ? 1 : 0);
          DP_LAUNCH_CHECK();
          return DOGPUP_OK;
        }
      }
      k_spmv_halo<NSP, MAT, EPI><<<vgrid(), 256, 0, st>>>(rp, m0, m1, x, y, w, coef, coef2, xacc, p1, p2, W, H,
                                                          ga, S->halo_ilv ? 1 : 0);
      DP_LAUNCH_CHECK();
    }
    return DOGPUP_OK;
  }
  template <int EPI>
  int spmvA(const T* x, T* y, const T* w, const int* ga) {
    pbegin();
    if (halo(0))
      DP_TRY((launch_halo<MAT_A, EPI>(0, S->rA, a0(), a1(), x, y, w, ga, nullptr, nullptr, nullptr)));
    else
      DP_TRY((launch_u<MAT_A, EPI>(S->unroll_a, S->rA, S->cA, a0(), a1(), x, y, w, ga, nullptr, nullptr, nullptr)));
    pend(1, rows(ga != nullptr) * rowbytes() * ((EPI == EPI_STORE || w == x) ? 2 : 3) +
                (double)S->nnzA * (2 * sizeof(T) + 4) + 4.0 * (S->nn + 1));
    return DOGPUP_OK;
  }
  template <int EPI>
  int spmvG(bool tr, const T* x, T* y, const T* w, const int* ga, const double* coef = nullptr,
            const double* coef2 = nullptr, T* xacc = nullptr) {
    pbegin();
    if (halo(tr ? 2 : 1))
      DP_TRY((launch_halo<MAT_G, EPI>(tr ? 2 : 1, tr ? S->rGT : S->rG, gvals(tr), nullptr, x, y, w, ga, coef, coef2,
                                      xacc)));
    else
      DP_TRY((launch_u<MAT_G, EPI>(S->unroll_g, tr ? S->rGT : S->rG, tr ? S->cGT : S->cG, gvals(tr), nullptr, x, y,
                                   w, ga, coef, coef2, xacc)));
    const int streams = (EPI == EPI_STORE || EPI == EPI_DOT2 || EPI == EPI_STORE_SELF || EPI == EPI_STORE_SELF2) ? 2
                        : (EPI == EPI_AXPBY_X ? 5 : 3);
    // G values: one copy per group, nnzG/nn entries per row on average
    pend(0, rows(ga != nullptr) * (rowbytes() * streams + (double)S->nnzG / S->nn * (2 * sizeof(T) + 4)) +
                4.0 * (S->nn + 1));
    return DOGPUP_OK;
  }

  int wait_count(int* out) {
    DP_HIP(hipMemcpyAsync(S->h_counter, bs.counter, 3 * sizeof(int), hipMemcpyDeviceToHost, st));
    DP_HIP(hipStreamSynchronize(st));
    *out = S->h_counter[0];
    S->active_groups = S->h_counter[1];
    S->active_rows = S->h_counter[2];
    return DOGPUP_OK;
  }
  int reset_count() {
    k_zero_ints<<<1, 64, 0, st>>>(bs.counter, 3);
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
  // bytes of one vector row (NSP columns) and the rows a pass touches: those
  // of the active groups when groups are skipped (polled at the last
  // convergence check), else all of the chunk's
  double rowbytes() const { return (double)NSP * 2 * sizeof(T); }
  double rows(bool skip) const { return skip ? (double)S->active_rows : (double)nodes; }

  // One BiCGSTAB iteration over every active group (no host synchronisation,
  // no memset nodes: it is captured into a HIP graph when profiling is off).
  int body(Vecs<T> V, int maxit) {
    const int* ga = bs.grp_active;
    const int nsp = NSP;
    DP_TRY(spmvG<EPI_STORE>(false, V.p, V.u, nullptr, ga));      // u = G p
    DP_TRY(spmvG<EPI_STORE>(true, V.u, V.h, nullptr, ga));       // h = G^T u
    DP_TRY(spmvA<EPI_STORE_DOT>(V.h, V.v, V.rhat, ga));          // v = A h, rhat.v
    k_sc_alpha<<<gcount, kScalThreads, 0, st>>>(bs, p1, nsp, W);
    DP_LAUNCH_CHECK();
    pbegin();
    k_axpy_s<T, NSP><<<vgrid(), 256, 0, st>>>(V.r, V.v, V.s, bs.alpha, p1, W, ga);
    DP_LAUNCH_CHECK();
    pend(2, rows(true) * rowbytes() * 3);
    k_sc_s<<<gcount, kScalThreads, 0, st>>>(bs, p1, nsp, W);
    DP_LAUNCH_CHECK();
    DP_TRY(spmvG<EPI_STORE>(false, V.s, V.u, nullptr, ga));      // u = G s
    DP_TRY(spmvG<EPI_STORE>(true, V.u, V.z, nullptr, ga));       // z = G^T G s
    DP_TRY(spmvA<EPI_STORE>(V.z, V.t, nullptr, ga));             // t = A z
    DP_TRY(spmvG<EPI_DOT2>(false, V.t, nullptr, V.u, ga));       // (Gt.Gs), (Gt.Gt)
    k_sc_omega<<<gcount, kScalThreads, 0, st>>>(bs, p1, p2, nsp, W);
    DP_LAUNCH_CHECK();
    pbegin();
    k_update_xr<T, NSP><<<vgrid(), 256, 0, st>>>(V.x, V.h, V.z, V.r, V.s, V.t, V.rhat, bs.alpha, bs.omega,
                                                bs.status, p1, p2, W, ga);
    DP_LAUNCH_CHECK();
    pend(2, rows(true) * rowbytes() * 8);
    DP_TRY(reset_count());
    const double rows_before = rows(true);  // groups still active before this update
    k_sc_beta<<<gcount, kScalThreads, 0, st>>>(bs, p1, p2, nsp, W, maxit);
    DP_LAUNCH_CHECK();
    pbegin();
    k_update_p<T, NSP><<<vgrid(), 256, 0, st>>>(V.p, V.r, V.v, bs.beta, bs.omega, W, ga);
    DP_LAUNCH_CHECK();
    pend(2, rows_before * rowbytes() * 4);
    return DOGPUP_OK;
  }

  // One COCG iteration over every active group: q = A p (p.q), alpha,
  // r -= alpha q (|r|^2), u = G r (u.u), convergence and beta, then
  // p = G^T u + beta p with x += alpha p (old p) in the same pass.  Three
  // operator applications and 12 vector streams per iteration (BiCGSTAB:
  // seven and 30).  The G^T pass runs for the groups that were active when
  // the iteration began (grp_prev), so a group that converges in this
  // iteration still receives its last x update.
  // Inner solve of the refinement form: COCG on the preconditioned
  // residual u = G r, updated as u -= alpha G q in the G pass itself, so r
  // is neither updated nor read again (3 passes and 10 vector streams per
  // iteration).  The column stops on |u|/|u_0|; the outer FP64 residual
  // check decides convergence.
  bool rfree() const { return std::is_same<T, float>::value && S->mode == DOGPUP_MODE_C64_REFINE && S->cg_rfree; }
  int body_cocg_rfree(Vecs<T> V, int maxit) {
    const int* ga = bs.grp_active;
    const int nsp = NSP;
    DP_TRY(spmvA<EPI_STORE_DOT>(V.p, V.v, V.p, ga));                     // q = A p, p.q
    k_cg_alpha<<<gcount, kScalThreads, 0, st>>>(bs, p1, nsp, W);
    DP_LAUNCH_CHECK();
    DP_TRY(spmvG<EPI_AXPY_SELF>(false, V.v, V.u, nullptr, ga, bs.alpha));  // u -= alpha G q; u.u, |u|^2
    DP_TRY(reset_count());
    k_cg_beta<<<gcount, kScalThreads, 0, st>>>(bs, p1, p2, nsp, W, maxit);
    DP_LAUNCH_CHECK();
    DP_TRY(spmvG<EPI_AXPBY_X>(true, V.u, V.p, nullptr, bs.grp_prev, bs.beta, bs.alpha, V.x));
    return DOGPUP_OK;
  }

  int body_cocg(Vecs<T> V, int maxit) {
    if (rfree()) return body_cocg_rfree(V, maxit);
    const int* ga = bs.grp_active;
    const int nsp = NSP;
    DP_TRY(spmvA<EPI_STORE_DOT>(V.p, V.v, V.p, ga));             // q = A p, p.q
    k_cg_alpha<<<gcount, kScalThreads, 0, st>>>(bs, p1, nsp, W);
    DP_LAUNCH_CHECK();
    pbegin();
    k_cg_update<T, NSP><<<vgrid(), 256, 0, st>>>(V.r, V.v, bs.alpha, p2, W, ga);
    DP_LAUNCH_CHECK();
    pend(2, rows(true) * rowbytes() * 3);
    DP_TRY(spmvG<EPI_STORE_SELF>(false, V.r, V.u, nullptr, ga));  // u = G r, u.u
    DP_TRY(reset_count());
    k_cg_beta<<<gcount, kScalThreads, 0, st>>>(bs, p1, p2, nsp, W, maxit);
    DP_LAUNCH_CHECK();
    // p = G^T u + beta p, x += alpha p
    DP_TRY(spmvG<EPI_AXPBY_X>(true, V.u, V.p, nullptr, bs.grp_prev, bs.beta, bs.alpha, V.x));
    return DOGPUP_OK;
  }

  // COCG start from r (state initialised, x = 0): u = G r, rho = u.u,
  // p = G^T u (stored: p's previous contents are not read).
  int cocg_start(Vecs<T> V) {
    const int* ga = bs.grp_active;
    if (rfree())
      DP_TRY(spmvG<EPI_STORE_SELF2>(false, V.r, V.u, nullptr, ga));
    else
      DP_TRY(spmvG<EPI_STORE_SELF>(false, V.r, V.u, nullptr, ga));
    DP_TRY(reset_count());
    k_cg_start<<<gcount, kScalThreads, 0, st>>>(bs, p1, rfree() ? p2 : nullptr, NSP, W);
    DP_LAUNCH_CHECK();
    DP_TRY(spmvG<EPI_STORE>(true, V.u, V.p, nullptr, ga));    // p = G^T u
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
      k_init<T, NSP><<<vgrid(), 256, 0, st>>>(b, V.x, V.r, V.rhat, V.p, p1, p2, W, nullptr);
      DP_LAUNCH_CHECK();
      DP_TRY(reset_count());
      k_sc_init<<<gcount, kScalThreads, 0, st>>>(bs, p1, p2, mask, nsp, W);
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
      const long long key = ((long long)phase << 52) ^ ((long long)cidx << 32) ^ ((long long)maxit << 3) ^
                            ((long long)S->method << 2) ^ (sizeof(T) == 4 ? 1 : 2);
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
// The refinement form's COCG (cocg_only) needs x, r, p, v and u alone.
template <typename T>
static Vecs<T> carve(const Buf& ws, size_t vbytes, bool own_x, bool with_w, bool cocg_only = false) {
  Vecs<T> v{};
  size_t off = 0;
  auto take = [&]() { T* p = ws.as<T>(off); off += align256(vbytes); return p; };
  v.x = own_x ? take() : nullptr;
  v.r = take();
  v.p = take();
  v.v = take();
  v.u = take();
  if (!cocg_only) {
    v.rhat = take(); v.h = take(); v.s = take(); v.z = take(); v.t = take();
  }
  v.w = with_w ? take() : nullptr;
  return v;
}
static bool refine_cocg(int mode, int method) { return mode == DOGPUP_MODE_C64_REFINE && method == DOGPUP_METHOD_COCG; }
static int nvec_for(int mode, int method) {
  if (refine_cocg(mode, method)) return 5;
  return mode == DOGPUP_MODE_C64_REFINE ? 10 : (mode == DOGPUP_MODE_C128_PLAIN ? 10 : 9);
}

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

// a[i] = t[i] where the per-RHS tolerance t is set (> 0), else v
__global__ void k_fill_tol(double* __restrict__ a, const double* __restrict__ t, double v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = (t && t[i] > 0.0) ? t[i] : v;
}

// Per-RHS tolerances of a chunk starting at group g0 (null: none set).
static const double* rhs_tol_at(const dogpup_solver* S, int g0) {
  return S->have_rhs_tol ? S->rhs_tol.as<double>() + (size_t)g0 * S->nsp : nullptr;
}

// Output index (freq*ncols + col within its mesh, meshes one after the
// other) of column c of group g, or -1 for padding.
static int out_index(const dogpup_solver* S, int g, int c) {
  if (c >= S->h_gcols[g]) return -1;
  const int m = S->h_gmesh[g];
  return S->m_rhs0[m] + S->h_gf[g] * S->m_ncols[m] + S->h_gcb[g] * S->nsp + c;
}

// Copy per-RHS results of one chunk into the host arrays.
static int export_results(dogpup_solver* S, hipStream_t st, int g0, int gcount, const int* d_iters,
                          const double* d_relres, const int* d_status, int* h_iters, double* h_relres,
                          int* h_status) {
  const int nrhs = gcount * S->nsp;
  std::vector<int> it(nrhs), stt(nrhs);
  std::vector<double> rr(nrhs);
  DP_HIP(hipMemcpyAsync(it.data(), d_iters, nrhs * 4, hipMemcpyDeviceToHost, st));
  DP_HIP(hipMemcpyAsync(stt.data(), d_status, nrhs * 4, hipMemcpyDeviceToHost, st));
  DP_HIP(hipMemcpyAsync(rr.data(), d_relres, nrhs * 8, hipMemcpyDeviceToHost, st));
  DP_HIP(hipStreamSynchronize(st));
  for (int g = 0; g < gcount; ++g)
    for (int c = 0; c < S->nsp; ++c) {
      const int o = out_index(S, g0 + g, c), i = g * S->nsp + c;
      if (o < 0) continue;
      int code = stt[i];
      if (code < 0) code = DOGPUP_RHS_MAXIT;
      if (h_iters) h_iters[o] = it[i];
      if (h_relres) h_relres[o] = rr[i];
      if (h_status) h_status[o] = code;
    }
  return DOGPUP_OK;
}

template <int NSP>
static int permute_block(dogpup_solver* S, hipStream_t st, const Work& W, const double* src, double* dst,
                         int dir) {
  k_permute_rows<NSP><<<dim3((unsigned)W.nw), 256, 0, st>>>(src, dst, S->order.as<int>(), dir, W);
  DP_LAUNCH_CHECK();
  return DOGPUP_OK;
}

template <typename T, int NSP>
static Chunk<T, NSP> make_chunk(dogpup_solver* S, hipStream_t st, int c) {
  Chunk<T, NSP> C;
  C.S = S;
  C.st = st;
  C.cidx = c;
  const int g0 = S->chunk_g[c], g1 = S->chunk_g[c + 1];
  C.gcount = g1 - g0;
  C.W = chunk_work(S, c);
  C.bs = carve_state(S->state, C.gcount * NSP, C.gcount);
  C.p1 = S->part1.as<double>();
  C.p2 = S->part2.as<double>();
  C.nodes = group_nodes(S, g0, g1);
  return C;
}

template <int NSP>
static int solve_c128(dogpup_solver* S, hipStream_t st, const double* b, double* x, double tol,
                      int maxit, int* h_iters, double* h_relres, int* h_status) {
  for (int c = 0; c + 1 < (int)S->chunk_g.size(); ++c) {
    Chunk<double, NSP> C = make_chunk<double, NSP>(S, st, c);
    const int gc = C.gcount, g0 = S->chunk_g[c];
    Vecs<double> V = carve<double>(S->ws, (size_t)C.nodes * NSP * 16, false, false);
    const size_t goff = (size_t)C.W.vb0 * NSP * 2;
    const bool ident = S->ident_solve;
    if (!ident) DP_TRY(permute_block<NSP>(S, st, C.W, b + goff, S->bperm.as<double>(), 0));
    V.x = ident ? x + goff : S->xperm.as<double>();
    const double* bc = ident ? b + goff : S->bperm.as<double>();
    k_fill_tol<<<div_up(gc * NSP, 256), 256, 0, st>>>(C.bs.tol, rhs_tol_at(S, g0), tol, gc * NSP);
    DP_LAUNCH_CHECK();
    int its = 0;
    DP_TRY(C.run_fused(V, bc, maxit, nullptr, &its));
    S->stats[0] = std::max(S->stats[0], (double)its);
    // true residual for reporting
    k_residual64<NSP><<<C.vgrid(), 256, 0, st>>>(S->rA, S->cA, S->a64pre.as<double>(), S->a64pim.as<double>(),
                                                 bc, nullptr, V.x, nullptr, nullptr, nullptr, nullptr, nullptr, 0,
                                                 C.p1, C.p2, C.W, S->halo_ilv ? 1 : 0);
    DP_LAUNCH_CHECK();
    const int g1 = S->chunk_g[c + 1];
    const int nparts = (g1 < S->ngroups ? S->h_gwbeg[g1] : S->nparts) - C.W.pw0;
    std::vector<double> part((size_t)nparts * NSP * 2);
    DP_HIP(hipMemcpyAsync(part.data(), C.p1, part.size() * 8, hipMemcpyDeviceToHost, st));
    std::vector<double> bn(gc * NSP);
    DP_HIP(hipMemcpyAsync(bn.data(), C.bs.bnorm, bn.size() * 8, hipMemcpyDeviceToHost, st));
    if (!ident) DP_TRY(permute_block<NSP>(S, st, C.W, V.x, x + goff, 1));
    DP_TRY(export_results(S, st, g0, gc, C.bs.iters, C.bs.relres, C.bs.status, h_iters, h_relres, h_status));
    // relres = |b - A x| / |b| from the residual pass's partials
    for (int g = 0; g < gc; ++g) {
      const int wb = S->h_gwbeg[g0 + g] - C.W.pw0, wc = S->h_gwcnt[g0 + g];
      int mx = 0;
      for (int cc = 0; cc < NSP; ++cc) {
        const int o = out_index(S, g0 + g, cc);
        if (o < 0) continue;
        double s2 = 0.0;
        for (int k = 0; k < wc; ++k) s2 += part[2 * ((size_t)(wb + k) * NSP + cc)];
        const int i = g * NSP + cc;
        if (h_relres) h_relres[o] = bn[i] > 0 ? std::sqrt(s2) / bn[i] : 0.0;
        if (h_iters) mx = std::max(mx, h_iters[o]);
      }
      S->stats[1] += mx;
    }
  }
  return DOGPUP_OK;
}

static int build_phases(dogpup_solver* S);

// a[i] = v for i < n
__global__ void k_fill_int(int* __restrict__ a, int v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// ------------------------------------------- Galerkin guesses, host side --
static_assert(kRbTerms <= kRbWeightsMax, "window size");

// Host data uploaded as one block: typed arrays appended at 256-byte-aligned
// offsets.
struct RbUpload {
  std::vector<char> bytes;
  template <typename V>
  size_t add(const std::vector<V>& v) {
    const size_t off = (bytes.size() + 255) & ~size_t(255);
    bytes.resize(off + std::max<size_t>(v.size() * sizeof(V), 1));
    if (!v.empty()) std::memcpy(bytes.data() + off, v.data(), v.size() * sizeof(V));
    return off;
  }
};

// The plan of every chunk's Galerkin guesses (dogpup_solver::RbPhase, one
// per phase), built with the phases: which phases' solutions get M x at the
// start of a phase, the Gram pairs its windows still need (each pair once
// over the solve, a new solution on the M x side where there is one), the
// per-target tables of the weight kernel and the guess launches.  It
// depends on the schedule only, so the solve issues the launches without
// waiting for any result.  A chunk whose windows do not fit the LDS staging
// keeps the Lagrange guesses (chunk_rb = 0).
static int upload(Buf& b, const void* h, size_t bytes);
static int rb_build_plan(dogpup_solver* S) {
  S->rb_phase.clear();
  const int nsp = S->nsp, ng = S->ngroups, nt = S->guess_nterm;
  const int PQ = 256 / nsp;
  const int kMaxPairs = PQ * kRbPairsPerThread, kMaxTgt = PQ * kRbTgtPerThread;
  RbUpload up;
  long long npair = 0;
  int max_tgt = 0;
  for (size_t c = 0; c < S->chunk_phases.size(); ++c) {
    std::vector<dogpup_solver::RbPhase> plan(S->chunk_phases[c].size());
    if (c >= S->chunk_rb.size() || !S->chunk_rb[c] || nt != kRbTerms) {
      S->rb_phase.push_back(plan);
      continue;
    }
    const auto& phases = S->chunk_phases[c];
    const int g0 = S->chunk_g[c];
    const long long vb0 = S->h_gvb[g0];
    auto groups_of = [&](const dogpup_solver::Phase& ph) {
      std::vector<int> gs;
      for (int w = ph.w0; w < ph.w0 + ph.nw; ++w)
        if (gs.empty() || gs.back() != S->h_pwgrp[w]) gs.push_back(S->h_pwgrp[w]);
      return gs;
    };
    auto family = [&](int g) { return (long long)S->h_gmesh[g] * 65536 + S->h_gcb[g]; };
    auto off_of = [&](int g) { return S->h_gvb[g] - vb0; };
    auto scale = [&](int g) { return S->h_fscale[S->m_fs0[S->h_gmesh[g]] + S->h_gf[g]]; };
    auto spec = [&](int g, double* re, double* im) {
      *re = 1.0;
      *im = 0.0;
      if (S->h_spec.empty()) return;
      const int f = S->m_fs0[S->h_gmesh[g]] + S->h_gf[g];
      *re = S->h_spec[2 * f];
      *im = S->h_spec[2 * f + 1];
    };
    std::vector<char> mx_done(phases.size(), 0);
    // computed pairs: (lo, hi) -> (pair id, computed as x_hi^H M x_lo)
    std::map<long long, std::pair<long long, bool>> pid;
    bool ok = true;
    for (size_t ip = 1; ip < phases.size() && ok; ++ip) {
      if (!phases[ip].guess) continue;
      dogpup_solver::RbPhase& P = plan[ip];
      const std::vector<int> tgt = groups_of(phases[ip]);
      std::vector<char> is_new(ng, 0);
      std::vector<int> glist;
      for (size_t q = 0; q < ip; ++q) {
        if (mx_done[q]) continue;
        mx_done[q] = 1;
        P.mx_phases.push_back((int)q);
        for (int g : groups_of(phases[q])) {
          is_new[g] = 1;
          glist.push_back(g - g0);
        }
      }
      P.o_glist = up.add(glist);
      P.nglist = (int)glist.size();
      std::map<long long, std::vector<int>> fam_tgt, fam_src;
      std::map<long long, std::vector<std::pair<int, int>>> fam_pairs;
      for (int t : tgt) {
        const int* w = &S->h_gsrc[(size_t)t * nt];
        const long long fk = family(t);
        fam_tgt[fk].push_back(t);
        for (int i = 0; i < nt && w[i] >= 0; ++i) {
          fam_src[fk].push_back(w[i]);
          for (int j = i; j < nt && w[j] >= 0; ++j) {
            const int lo = std::min(w[i], w[j]), hi = std::max(w[i], w[j]);
            if (pid.count((long long)lo * ng + hi)) continue;
            if (!is_new[hi] && is_new[lo]) fam_pairs[fk].push_back({hi, lo});
            else fam_pairs[fk].push_back({lo, hi});
          }
        }
      }
      for (auto& kv : fam_src) {
        std::sort(kv.second.begin(), kv.second.end());
        kv.second.erase(std::unique(kv.second.begin(), kv.second.end()), kv.second.end());
        ok &= kv.second.size() * nsp * 16 <= (size_t)kRbLds;
      }
      // Gram launches, at most kMaxPairs pairs each
      for (auto& kv : fam_pairs) {
        auto& pr = kv.second;
        std::sort(pr.begin(), pr.end());
        pr.erase(std::unique(pr.begin(), pr.end()), pr.end());
        for (size_t i0 = 0; i0 < pr.size() && ok; i0 += kMaxPairs) {
          const size_t i1 = std::min(pr.size(), i0 + (size_t)kMaxPairs);
          std::vector<int> la, lb;
          for (size_t i = i0; i < i1; ++i) {
            la.push_back(pr[i].first);
            lb.push_back(pr[i].second);
          }
          std::sort(la.begin(), la.end());
          la.erase(std::unique(la.begin(), la.end()), la.end());
          std::sort(lb.begin(), lb.end());
          lb.erase(std::unique(lb.begin(), lb.end()), lb.end());
          const size_t row_bytes = (la.size() + lb.size()) * nsp * 16;
          if (row_bytes > (size_t)kRbLds) {
            ok = false;
            break;
          }
          dogpup_solver::RbGram L;
          L.R = (int)std::min<size_t>(8, kRbLds / row_bytes);
          std::vector<long long> oa, ob;
          for (int g : la) oa.push_back(off_of(g));
          for (int g : lb) ob.push_back(off_of(g));
          std::vector<int> pia, pib;
          for (size_t i = i0; i < i1; ++i) {
            pia.push_back((int)(std::lower_bound(la.begin(), la.end(), pr[i].first) - la.begin()));
            pib.push_back((int)(std::lower_bound(lb.begin(), lb.end(), pr[i].second) - lb.begin()));
            const int lo = std::min(pr[i].first, pr[i].second), hi = std::max(pr[i].first, pr[i].second);
            pid[(long long)lo * ng + hi] = {npair + (long long)(i - i0), pr[i].first > pr[i].second};
          }
          L.na = (int)la.size();
          L.nb = (int)lb.size();
          L.np = (int)(i1 - i0);
          L.nrows = S->h_grows[pr[i0].first];
          L.o_offa = up.add(oa);
          L.o_offb = up.add(ob);
          L.o_pia = up.add(pia);
          L.o_pib = up.add(pib);
          L.out = npair;
          npair += L.np;
          P.gram.push_back(L);
        }
      }
      if (!ok) break;
      // the weight kernel's table per target: window size, columns, pair
      // codes (2 * pair id + conjugate) of the window's Gram entries, the
      // window's groups (chunk-local, for x^H b), spectra and scales
      std::vector<int> tinfo;
      std::vector<double> tspec;
      std::vector<int> order_t;   // targets in the order of the guess launches
      for (auto& kv : fam_tgt)
        for (int t : kv.second) order_t.push_back(t);
      for (int t : order_t) {
        const int* win = &S->h_gsrc[(size_t)t * nt];
        int n = 0;
        while (n < nt && win[n] >= 0) ++n;
        tinfo.push_back(n);
        tinfo.push_back(S->h_gcols[t]);
        for (int a = 0; a < kRbTerms; ++a)
          for (int b = 0; b < kRbTerms; ++b) {
            int code = -1;
            if (a < n && b < n) {
              const int lo = std::min(win[a], win[b]), hi = std::max(win[a], win[b]);
              const auto e = pid.find((long long)lo * ng + hi);
              if (e == pid.end()) {   // every window pair is planned above
                ok = false;
                break;
              }
              // stored value x_lo^H M x_hi, or its conjugate when computed
              // the other way round; entry (a, b) is x_win[a]^H M x_win[b]
              const bool conj = e->second.second != (win[a] > win[b]);
              code = (int)(2 * e->second.first + (conj ? 1 : 0));
            }
            tinfo.push_back(code);
          }
        for (int a = 0; a < kRbTerms; ++a) tinfo.push_back(a < n ? win[a] - g0 : 0);
        double re, im;
        for (int a = 0; a < kRbTerms; ++a) {
          if (a < n) spec(win[a], &re, &im);
          else re = im = 0.0;
          tspec.push_back(re);
          tspec.push_back(im);
        }
        for (int a = 0; a < kRbTerms; ++a) tspec.push_back(a < n ? scale(win[a]) : 0.0);
        spec(t, &re, &im);
        tspec.push_back(re);
        tspec.push_back(im);
        tspec.push_back(scale(t));
      }
      if (!ok) break;
      P.o_tinfo = up.add(tinfo);
      P.o_tspec = up.add(tspec);
      P.ntgt = (int)order_t.size();
      max_tgt = std::max(max_tgt, P.ntgt);
      // guess launches: at most kMaxTgt targets of one mesh and column block
      int tpos = 0;
      for (auto& kv : fam_tgt) {
        const std::vector<int>& src = fam_src[kv.first];
        const auto& tg = kv.second;
        for (size_t i0 = 0; i0 < tg.size(); i0 += kMaxTgt) {
          const size_t i1 = std::min(tg.size(), i0 + (size_t)kMaxTgt);
          std::vector<long long> oa, ot;
          for (int g : src) oa.push_back(off_of(g));
          std::vector<int> si;
          for (size_t i = i0; i < i1; ++i) {
            const int t = tg[i];
            const int* win = &S->h_gsrc[(size_t)t * nt];
            ot.push_back(off_of(t));
            for (int q = 0; q < kRbTerms; ++q)
              si.push_back(q < nt && win[q] >= 0
                               ? (int)(std::lower_bound(src.begin(), src.end(), win[q]) - src.begin())
                               : -1);
          }
          dogpup_solver::RbGuess L;
          L.na = (int)src.size();
          L.nt = (int)(i1 - i0);
          L.nrows = S->h_grows[tg[i0]];
          L.R = (int)std::min<size_t>(8, kRbLds / ((size_t)std::max(L.na, 1) * nsp * 16));
          L.o_offa = up.add(oa);
          L.o_offt = up.add(ot);
          L.o_src = up.add(si);
          L.w0 = tpos + (int)i0;
          P.guess.push_back(L);
        }
        tpos += (int)tg.size();
      }
      P.ok = true;
    }
    if (!ok) {
      S->chunk_rb[c] = 0;
      plan.assign(phases.size(), dogpup_solver::RbPhase{});
    }
    S->rb_phase.push_back(plan);
  }
  if (up.bytes.empty()) return DOGPUP_OK;
  DP_TRY(upload(S->rb_plan, up.bytes.data(), up.bytes.size()));
  const size_t max_groups = (size_t)ng;
  DP_TRY(S->rb_cdot.ensure(max_groups * nsp * 16));
  DP_TRY(S->rb_gram.ensure((size_t)std::max<long long>(npair, 1) * nsp * 16));
  DP_TRY(S->rb_w.ensure((size_t)std::max(max_tgt, 1) * kRbTerms * nsp * 16));
  DP_TRY(S->rb_part.ensure((size_t)kRbBlocks * kMaxPairs * nsp * 16));
  return DOGPUP_OK;
}

// The launches of phase ip's Galerkin guesses (rb_build_plan): M x and
// x^H b of the planned phases' solutions, the Gram entries, the weights on
// the device, and the guesses, all on the stream without a host wait.
// x64 (xc) holds every earlier phase's converged solutions; bperm receives
// M x at the same offsets.
template <int NSP>
static int rb_run_phase(dogpup_solver* S, hipStream_t st, int c, const Work& Wfull, double* p1,
                        const std::vector<dogpup_solver::Phase>& phases, size_t ip, const double* bc, double* xc) {
  const dogpup_solver::RbPhase& P = S->rb_phase[c][ip];
  const char* dm = S->rb_plan.as<char>();
  double* cdot = S->rb_cdot.as<double>();
  for (int q : P.mx_phases) {
    const dogpup_solver::Phase& ph = phases[q];
    Work Wp = Wfull;
    Wp.wgrp = S->t_pwgrp.as<int>();
    Wp.wtile = S->t_pwtile.as<int>();
    Wp.wpart = S->t_pwpart.as<int>();
    Wp.w0 = ph.w0;
    Wp.nw = ph.nw;
    k_rb_mx<NSP><<<dim3((unsigned)ph.nw), 256, 0, st>>>(S->rA, S->cA, S->a64pim.as<double>(), xc,
                                                        S->bperm.as<double>(), bc, S->order.as<int>(), p1, Wp,
                                                        S->halo_ilv ? 1 : 0);
    DP_LAUNCH_CHECK();
  }
  if (P.nglist > 0) {
    k_rb_csum<<<dim3((unsigned)P.nglist), kScalThreads, 0, st>>>(
        p1, cdot, reinterpret_cast<const int*>(dm + P.o_glist), NSP, Wfull);
    DP_LAUNCH_CHECK();
  }
  for (const auto& L : P.gram) {
    const int ntile = (L.nrows + L.R - 1) / L.R, nblk = std::min(kRbBlocks, ntile);
    const size_t lds = (size_t)(L.na + L.nb) * L.R * NSP * 16;
    k_rb_gram<NSP><<<dim3((unsigned)nblk), 256, lds, st>>>(
        xc, S->bperm.as<double>(), reinterpret_cast<const long long*>(dm + L.o_offa), L.na,
        reinterpret_cast<const long long*>(dm + L.o_offb), L.nb, reinterpret_cast<const int*>(dm + L.o_pia),
        reinterpret_cast<const int*>(dm + L.o_pib), L.np, L.nrows, L.R, S->rb_part.as<double>());
    DP_LAUNCH_CHECK();
    const int n = L.np * NSP;
    k_rb_reduce<<<dim3((unsigned)div_up(n, 256)), 256, 0, st>>>(S->rb_part.as<double>(),
                                                               S->rb_gram.as<double>() + 2 * L.out * NSP, n, nblk);
    DP_LAUNCH_CHECK();
  }
  if (P.ntgt > 0) {
    k_rb_weights<<<dim3(div_up((long long)P.ntgt * NSP, 64)), 64, 0, st>>>(
        reinterpret_cast<const int*>(dm + P.o_tinfo), reinterpret_cast<const double*>(dm + P.o_tspec),
        S->rb_gram.as<double>(), cdot, S->rb_w.as<double>(), P.ntgt, NSP);
    DP_LAUNCH_CHECK();
  }
  for (const auto& L : P.guess) {
    const int ntile = (L.nrows + L.R - 1) / L.R, nblk = std::min(kRbBlocks, ntile);
    const size_t lds = (size_t)std::max(L.na, 1) * L.R * NSP * 16;
    k_rb_guess<NSP><<<dim3((unsigned)nblk), 256, lds, st>>>(
        xc, reinterpret_cast<const long long*>(dm + L.o_offa), L.na, reinterpret_cast<const long long*>(dm + L.o_offt),
        L.nt, reinterpret_cast<const int*>(dm + L.o_src), S->rb_w.as<double>() + 2 * (size_t)L.w0 * kRbTerms * NSP,
        L.nrows, L.R);
    DP_LAUNCH_CHECK();
  }
  return DOGPUP_OK;
}

template <int NSP>
static int solve_refine(dogpup_solver* S, hipStream_t st, const double* b, double* x, double tol,
                        int maxit, int* h_iters, double* h_relres, int* h_status) {
  const char* env_tol = getenv("DOGPUP_INNER_TOL");
  const double tol_in = env_tol ? atof(env_tol) : 1e-4;
  if (S->guess_nterm > 0 && S->phases_dirty) DP_TRY(build_phases(S));
  const bool phased = S->guess_nterm > 0;
  for (int c = 0; c + 1 < (int)S->chunk_g.size(); ++c) {
    Chunk<float, NSP> C = make_chunk<float, NSP>(S, st, c);
    const int gc = C.gcount, g0 = S->chunk_g[c];
    RState R = carve_rstate(S->rstate, gc * NSP);
    Vecs<float> V = carve<float>(S->ws, (size_t)C.nodes * NSP * 8, true, false, refine_cocg(S->mode, S->method));
    const size_t goff = (size_t)C.W.vb0 * NSP * 2;
    // b is read in the caller's order by every residual pass; x64 (solver
    // order) is first written by the first sweep's correction, or by the
    // guess of its phase
    const double* bc = b + goff;
    const bool ident = S->ident_solve;
    double* xc = ident ? x + goff : S->xperm.as<double>();
    const int* border = ident ? nullptr : S->order.as<int>();   // b's rows in the caller's order
    const bool cocg = S->method == DOGPUP_METHOD_COCG;
    const Work Wfull = C.W;
    std::vector<dogpup_solver::Phase> one{{0, Wfull.w0, Wfull.nw, false, C.nodes}};
    const std::vector<dogpup_solver::Phase>& phases = phased ? S->chunk_phases[c] : one;
    if (phased) {
      // groups of later phases take no part until their phase
      k_fill_int<<<div_up(gc * NSP, 256), 256, 0, st>>>(R.status, ST_SKIP, gc * NSP);
      DP_LAUNCH_CHECK();
      if (S->guess_self && !ident) DP_TRY(permute_block<NSP>(S, st, Wfull, x + goff, xc, 0));
    }
    for (size_t ip = 0; ip < phases.size(); ++ip) {
      const dogpup_solver::Phase& ph = phases[ip];
      C.W = Wfull;
      if (phased) {
        C.W.wgrp = S->t_pwgrp.as<int>();
        C.W.wtile = S->t_pwtile.as<int>();
        C.W.wpart = S->t_pwpart.as<int>();
        C.W.w0 = ph.w0;
        C.W.nw = ph.nw;
      }
      C.phase = (int)ip;
      const double e64 = (double)ph.nodes * NSP;  // complex elements of the phase
      if (ph.guess) {
        const bool rb_done = c < (int)S->rb_phase.size() && ip < S->rb_phase[c].size() && S->rb_phase[c][ip].ok;
        if (rb_done) DP_TRY(rb_run_phase<NSP>(S, st, c, Wfull, C.p1, phases, ip, bc, xc));
        else {
          k_guess<NSP><<<C.vgrid(), 256, 0, st>>>(xc, C.W, S->t_gsrc.as<int>(), S->t_gw.as<double>(), S->guess_nterm,
                                                 x, S->order.as<int>());
          DP_LAUNCH_CHECK();
        }
      }
      for (int sweep = 0;; ++sweep) {
        const bool from0 = sweep == 0 && !ph.guess;   // x64 = 0: the residual is b
        double* p3 = (sweep == 0 && ph.guess) ? S->part3.as<double>() : nullptr;
        C.pbegin();
        // COCG starts from r alone (p = G^T G r is written by its start), so
        // only BiCGSTAB needs p and rhat
        k_residual64<NSP><<<C.vgrid(), 256, 0, st>>>(
            S->rA, S->cA, S->a64pre.as<double>(), S->a64pim.as<double>(), bc, border,
            from0 ? nullptr : xc, V.r, cocg ? nullptr : V.rhat, cocg ? nullptr : V.p, V.x,
            sweep == 0 ? nullptr : R.status, 1, C.p1, C.p2, C.W, S->halo_ilv ? 1 : 0, p3);
        DP_LAUNCH_CHECK();
        C.pend(3, e64 * (16 + (from0 ? 0 : 16) + (cocg ? 2 : 4) * 8) + (from0 ? 0.0 : (double)S->nnzA * 20) +
                      4.0 * (S->nn + 1));
        DP_TRY(C.reset_count());
        k_rf_check<<<gc, kScalThreads, 0, st>>>(R, C.bs, C.p1, NSP, C.W, sweep, tol, tol_in, maxit,
                                                 rhs_tol_at(S, g0), p3, phased ? S->t_glevel.as<int>() : nullptr,
                                                 ph.level);
        DP_LAUNCH_CHECK();
        if (sweep == 0 && ph.guess) {
          // zero right-hand sides (and padding columns) have x = 0
          k_zero_cols<NSP><<<C.vgrid(), 256, 0, st>>>(xc, R.status, C.W);
          DP_LAUNCH_CHECK();
        }
        int active = 0;
        DP_TRY(C.wait_count(&active));
        if (active == 0) {
          // every column met the tolerance at its first iterate: x = 0 for
          // zero right-hand sides, the guess otherwise
          if (from0) {
            k_guess<NSP><<<C.vgrid(), 256, 0, st>>>(xc, C.W, nullptr, nullptr, 0, nullptr, nullptr);
            DP_LAUNCH_CHECK();
          }
          break;
        }
        const double sweep_rows = (double)S->active_rows;
        // inner state from the residual pass: bnorm = |r64|, rho = r32.r32
        k_sc_init<<<gc, kScalThreads, 0, st>>>(C.bs, C.p1, C.p2, R.status, NSP, C.W);
        DP_LAUNCH_CHECK();
        int its = 0;
        DP_TRY(C.run_fused(V, nullptr, maxit, R.status, &its));
        C.pbegin();
        k_add_correction<NSP><<<C.vgrid(), 256, 0, st>>>(xc, V.x, C.bs.status, from0, C.W);
        DP_LAUNCH_CHECK();
        C.pend(3, (from0 ? e64 * 16 : 0.0) + sweep_rows * NSP * (from0 ? 8 : 32 + 8));
        k_rf_accum<<<gc, kScalThreads, 0, st>>>(R, C.bs, NSP);
        DP_LAUNCH_CHECK();
        S->stats[0] = std::max(S->stats[0], (double)its);
        S->stats[2] += 1;
        S->stats[3] += its;
        if (sweep > 64) break;
      }
    }
    C.W = Wfull;
    if (!ident) DP_TRY(permute_block<NSP>(S, st, C.W, xc, x + goff, 1));
    DP_TRY(export_results(S, st, g0, gc, R.iters, R.relres, R.status, h_iters, h_relres, h_status));
  }
  return DOGPUP_OK;
}

template <int NSP>
static int solve_plain(dogpup_solver* S, hipStream_t st, const double* b, double* x, double tol,
                       int maxit, int* h_iters, double* h_relres, int* h_status) {
  for (int c = 0; c + 1 < (int)S->chunk_g.size(); ++c) {
    Chunk<double, NSP> C = make_chunk<double, NSP>(S, st, c);
    const int gc = C.gcount, g0 = S->chunk_g[c];
    const Work& W = C.W;
    const size_t vbytes = (size_t)C.nodes * NSP * 16;
    Vecs<double> V = carve<double>(S->ws, vbytes, false, true);
    const size_t goff = (size_t)W.vb0 * NSP * 2;
    const bool ident = S->ident_solve;
    if (!ident) DP_TRY(permute_block<NSP>(S, st, W, b + goff, S->bperm.as<double>(), 0));
    V.x = ident ? x + goff : S->xperm.as<double>();
    const double* bc = ident ? b + goff : S->bperm.as<double>();
    unsigned long long* dmax = reinterpret_cast<unsigned long long*>(C.bs.counter);
    const dim3 vg = C.vgrid();
    auto maxerr = [&](int op, const double* pa, double* out) -> int {
      DP_HIP(hipMemsetAsync(dmax, 0, 8, st));
      k_plain_scal<<<gc, kScalThreads, 0, st>>>(C.bs, op, pa, nullptr, NSP, W, dmax);
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
    k_init<double, NSP><<<vg, 256, 0, st>>>(bc, V.x, V.r, V.rhat, V.p, C.p1, C.p2, W, nullptr);
    DP_LAUNCH_CHECK();
    k_plain_scal<<<gc, kScalThreads, 0, st>>>(C.bs, 5, C.p1, C.p2, NSP, W, dmax);
    DP_LAUNCH_CHECK();
    int it = 0;
    for (it = 1; it <= maxit; ++it) {
      DP_TRY(C.template spmvG<EPI_STORE>(false, V.p, V.u, nullptr, nullptr));   // h = G p
      DP_TRY(C.template spmvG<EPI_STORE>(true, V.u, V.h, nullptr, nullptr));    // h = G^T h
      DP_TRY(C.template spmvA<EPI_STORE>(V.h, V.v, nullptr, nullptr));          // v = A h
      k_dot<double, NSP><<<vg, 256, 0, st>>>(V.v, V.rhat, 0, C.p1, W, nullptr);
      DP_LAUNCH_CHECK();
      k_plain_scal<<<gc, kScalThreads, 0, st>>>(C.bs, 0, C.p1, nullptr, NSP, W, dmax);
      DP_LAUNCH_CHECK();
      k_axpby<double, NSP><<<vg, 256, 0, st>>>(V.u, V.x, V.h, C.bs.alpha, 0, W, nullptr);  // hx = x + h/alph
      DP_LAUNCH_CHECK();
      k_axpby<double, NSP><<<vg, 256, 0, st>>>(V.s, V.r, V.v, C.bs.alpha, 1, W, nullptr);  // s = r - v/alph
      DP_LAUNCH_CHECK();
      k_dot<double, NSP><<<vg, 256, 0, st>>>(V.s, nullptr, 1, C.p1, W, nullptr);
      DP_LAUNCH_CHECK();
      double err = 0;
      DP_TRY(maxerr(1, C.p1, &err));
      if (err < tol) {
        DP_HIP(hipMemcpyAsync(V.x, V.u, vbytes, hipMemcpyDeviceToDevice, st));
        break;
      }
      std::swap(V.h, V.u);  // h now holds x + h/alph
      DP_TRY(C.template spmvG<EPI_STORE>(false, V.s, V.u, nullptr, nullptr));   // Gs
      DP_TRY(C.template spmvG<EPI_STORE>(true, V.u, V.z, nullptr, nullptr));    // z
      DP_TRY(C.template spmvA<EPI_STORE>(V.z, V.t, nullptr, nullptr));          // t
      // Gtt = G t is materialised, as the reference does
      DP_TRY(C.template spmvG<EPI_STORE>(false, V.t, V.w, nullptr, nullptr));
      k_dot<double, NSP><<<vg, 256, 0, st>>>(V.w, V.u, 0, C.p1, W, nullptr);
      DP_LAUNCH_CHECK();
      k_dot<double, NSP><<<vg, 256, 0, st>>>(V.w, V.w, 0, C.p2, W, nullptr);
      DP_LAUNCH_CHECK();
      k_plain_scal<<<gc, kScalThreads, 0, st>>>(C.bs, 2, C.p1, C.p2, NSP, W, dmax);
      DP_LAUNCH_CHECK();
      k_axpby<double, NSP><<<vg, 256, 0, st>>>(V.x, V.h, V.z, C.bs.omega, 2, W, nullptr);  // x = h + om z
      DP_LAUNCH_CHECK();
      k_axpby<double, NSP><<<vg, 256, 0, st>>>(V.r, V.s, V.t, C.bs.omega, 3, W, nullptr);  // r = s - om t
      DP_LAUNCH_CHECK();
      k_axpby<double, NSP><<<vg, 256, 0, st>>>(V.p, V.p, V.v, C.bs.omega, 3, W, nullptr);  // p = p - om v
      DP_LAUNCH_CHECK();
      k_dot<double, NSP><<<vg, 256, 0, st>>>(V.r, nullptr, 1, C.p1, W, nullptr);
      DP_LAUNCH_CHECK();
      DP_TRY(maxerr(3, C.p1, &err));
      if (err < tol) break;
      k_dot<double, NSP><<<vg, 256, 0, st>>>(V.rhat, V.r, 0, C.p1, W, nullptr);
      DP_LAUNCH_CHECK();
      k_plain_scal<<<gc, kScalThreads, 0, st>>>(C.bs, 4, C.p1, nullptr, NSP, W, dmax);
      DP_LAUNCH_CHECK();
      k_axpby<double, NSP><<<vg, 256, 0, st>>>(V.p, V.r, V.p, C.bs.beta, 2, W, nullptr);  // p = r + beta p
      DP_LAUNCH_CHECK();
    }
    if (!ident) DP_TRY(permute_block<NSP>(S, st, W, V.x, x + goff, 1));
    S->stats[0] = std::max(S->stats[0], (double)std::min(it, maxit));
    S->stats[1] += (double)std::min(it, maxit) * gc;
    // status: converged if the lockstep loop met tol
    std::vector<double> rr(gc * NSP), bn(gc * NSP);
    DP_HIP(hipMemcpyAsync(rr.data(), C.bs.relres, rr.size() * 8, hipMemcpyDeviceToHost, st));
    DP_HIP(hipMemcpyAsync(bn.data(), C.bs.bnorm, bn.size() * 8, hipMemcpyDeviceToHost, st));
    DP_HIP(hipStreamSynchronize(st));
    for (int g = 0; g < gc; ++g)
      for (int cc = 0; cc < NSP; ++cc) {
        const int o = out_index(S, g0 + g, cc), i = g * NSP + cc;
        if (o < 0) continue;
        if (h_iters) h_iters[o] = std::min(it, maxit);
        if (h_relres) h_relres[o] = rr[i];
        if (h_status) h_status[o] = bn[i] == 0.0 ? DOGPUP_RHS_ZERO : (rr[i] < tol ? DOGPUP_RHS_CONVERGED
                                                                                  : DOGPUP_RHS_MAXIT);
      }
  }
  return DOGPUP_OK;
}

}  // namespace dogpup

namespace dogpup {

// Solver-layout transforms and point-source right-hand sides (one mesh)
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
    default: return F<64>::run(a...);
  }
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

// Workspace budget: DOGPUP_WORKSPACE_GB, else the smaller of 160 GB and half
// of the device memory free when the handle is made once the caller's
// complex128 b and x of this handle (io_bytes each) are set aside, whether or
// not they exist yet; larger batches run in group chunks.  It is computed
// once per handle, so re-tuning does not re-plan the chunks around memory the
// caller's allocator happens to hold at the time.
static double ws_budget_bytes(double io_bytes) {
  if (const char* e = getenv("DOGPUP_WORKSPACE_GB")) {
    const double gb = atof(e);
    if (gb > 0) return gb * 1073741824.0;
  }
  size_t free_b = 0, total_b = 0;
  double budget = 160.0 * 1073741824.0;
  if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b > 0)
    budget = std::min(budget, 0.5 * std::max(0.0, (double)free_b - 2.0 * io_bytes));
  return budget;
}

static int upload(Buf& b, const void* h, size_t bytes) {
  DP_TRY(b.ensure(bytes));
  DP_HIP(hipMemcpy(b.p, h, bytes, hipMemcpyHostToDevice));
  return DOGPUP_OK;
}

// Work list for the current tile size, the chunk plan for the workspace
// budget, and the buffers sized by them.
static int build_work(dogpup_solver* S) {
  const int ng = S->ngroups, tr = S->tile_rows;
  std::vector<int> wg, wt, wp;
  S->h_gibeg.assign(ng, 0);
  S->h_gwbeg.assign(ng, 0);
  S->h_gwcnt.assign(ng, 0);
  int parts = 0;
  for (int g = 0; g < ng; ++g) {
    const int nt = (S->h_grows[g] + tr - 1) / tr;
    const int ntp = (nt + 7) / 8 * 8, per = ntp / 8;
    S->h_gibeg[g] = (int)wg.size();
    S->h_gwbeg[g] = parts;
    S->h_gwcnt[g] = nt;
    // item k runs on XCD k % 8 and takes tile (k % 8) * per + k / 8
    for (int k = 0; k < ntp; ++k) {
      const int t = (k & 7) * per + (k >> 3);
      wg.push_back(g);
      wt.push_back(t < nt ? t : -1);
      wp.push_back(t < nt ? parts + t : 0);
    }
    parts += nt;
  }
  S->nwork = (int)wg.size();
  S->nparts = parts;
  S->h_wgrp = wg;
  S->h_wtile = wt;
  S->h_wpart = wp;
  DP_TRY(upload(S->t_wgrp, wg.data(), wg.size() * 4));
  DP_TRY(upload(S->t_wtile, wt.data(), wt.size() * 4));
  DP_TRY(upload(S->t_wpart, wp.data(), wp.size() * 4));
  DP_TRY(upload(S->t_gwbeg, S->h_gwbeg.data(), ng * 4));
  DP_TRY(upload(S->t_gwcnt, S->h_gwcnt.data(), ng * 4));
  // chunks: consecutive groups whose vectors fit the workspace budget
  const size_t csize = S->mode == DOGPUP_MODE_C64_REFINE ? 8 : 16;
  const double per_node = (double)S->nsp * (csize * nvec_for(S->mode, S->method) + 2.0 * 16);
  if (S->ws_budget <= 0) {
    double io_rows = 0;
    for (int g = 0; g < ng; ++g) io_rows += S->h_grows[g];
    S->ws_budget = ws_budget_bytes(io_rows * S->nsp * 16.0);
  }
  const double budget = S->ws_budget;
  S->chunk_g.assign(1, 0);
  double acc = 0;
  for (int g = 0; g < ng; ++g) {
    const double bg = per_node * S->h_grows[g];
    if (acc > 0 && acc + bg > budget) {
      S->chunk_g.push_back(g);
      acc = 0;
    }
    acc += bg;
  }
  S->chunk_g.push_back(ng);
  long long max_nodes = 0;
  int max_groups = 0, max_work = 0;
  for (size_t c = 0; c + 1 < S->chunk_g.size(); ++c) {
    const int g0 = S->chunk_g[c], g1 = S->chunk_g[c + 1];
    max_nodes = std::max(max_nodes, group_nodes(S, g0, g1));
    max_groups = std::max(max_groups, g1 - g0);
    max_work = std::max(max_work, (g1 < ng ? S->h_gwbeg[g1] : S->nparts) - S->h_gwbeg[g0]);
  }
  const size_t vbytes = align256((size_t)max_nodes * S->nsp * csize);
  DP_TRY(S->ws.ensure((size_t)nvec_for(S->mode, S->method) * vbytes));
  DP_TRY(S->bperm.ensure((size_t)max_nodes * S->nsp * 16));
  DP_TRY(S->xperm.ensure((size_t)max_nodes * S->nsp * 16));
  const size_t pbytes = (size_t)max_work * S->nsp * 16;
  DP_TRY(S->part1.ensure(pbytes));
  DP_TRY(S->part2.ensure(pbytes));
  DP_TRY(S->part3.ensure(pbytes));
  S->phases_dirty = true;
  DP_TRY(S->state.ensure(state_bytes(max_groups * S->nsp, max_groups)));
  DP_TRY(S->rstate.ensure(5 * align256((size_t)max_groups * S->nsp * 8)));
  S->clear_graphs();
  return DOGPUP_OK;
}

// Frequency continuation schedule of one mesh with nf frequencies at the
// scales fs[0..nf): frequency k is in level 0 when k % stride == 0 or k is
// the last, and one level later per halving of the spacing (stride 8: level
// 1 when k % 8 == 4, level 2 when k % 4 == 2 and level 3 otherwise).  A
// frequency of a later level starts from a guess built from the frequencies
// already solved nearest to it (ties to the lower index): those of earlier
// levels in its chunk, and every frequency of an earlier chunk.
//  * Galerkin guesses (the default, stride 16): the kRbTerms nearest, with
//    per-column weights from the small Galerkin system (k_rb_* above).  On
//    a 15,376-node slab with 64 frequencies (0-3.15 GHz) they took 241
//    COCG iterations over all frequencies to a 5e-8 residual, against 508
//    for the Lagrange guesses and 1,382 from x0 = 0.
//  * Lagrange guesses (stride 8): the kGuessTerms nearest, weighted by the
//    Lagrange interpolation in the frequency scale.  The diffusion fields
//    are smooth in the modulation frequency: on the 50 MHz grid of a 20 ns
//    period the cubic interpolation of the odd frequencies from their even
//    neighbours leaves a first residual of 1e-8..3e-6 of |b| on the
//    269k-node slab, against |b| itself from x0 = 0.  When the right-hand
//    side of frequency k is spec[k] times a vector that does not depend on
//    the frequency (point sources: the source spectrum, dMesh.m:128,196), it
//    is x / spec that is smooth (the spectrum of a delayed pulse turns by
//    0.2 rad per step), so a source j's term carries spec[k] / spec[j] (spec
//    NULL: flat).
// Frequencies with a zero spectrum are no source.
constexpr int kGuessTerms = 4;
constexpr int kMaxGuessTerms = 8;
static_assert(kRbTerms <= kMaxGuessTerms, "guess terms");
// Level of frequency index k: 0 for every stride-th and the last, then one
// level per halving of the spacing.  DOGPUP_CONT_STRIDE (a power of two)
// overrides the stride.
static int cont_stride(const dogpup_solver* S) {
  static const int env = [] {
    const char* e = getenv("DOGPUP_CONT_STRIDE");
    const int v = e ? atoi(e) : 0;
    return (v >= 2 && !(v & (v - 1))) ? v : 0;
  }();
  if (S->cont_stride > 0) return S->cont_stride;
  return env ? env : (S->cont_rb ? 16 : 8);
}
static int cont_level(int k, int nf, int st) {
  if (k % st == 0 || k == nf - 1) return 0;
  int L = 0, tz = 0;
  for (int v = st; v > 1; v >>= 1) ++L;
  while (!((k >> tz) & 1)) ++tz;
  return L - tz;
}
// Terms of frequency k: the nterm frequencies j with avail(j) nearest to k
// (distinct scales, nonzero spectrum), with the Lagrange weights at fs[k]
// times spec[k] / spec[j].  Returns the number of terms.
template <typename Avail>
static int cont_terms(const double* fs, const double* spec, int nf, int k, Avail avail, int nterm, int* sk,
                      double* wk) {
  auto sp = [&](int j) { return spec ? cpx<double>{spec[2 * j], spec[2 * j + 1]} : cpx<double>{1.0, 0.0}; };
  std::vector<int> cand;
  for (int j = 0; j < nf; ++j) {
    const cpx<double> sj = sp(j);
    if (j != k && avail(j) && (sj.re != 0.0 || sj.im != 0.0)) cand.push_back(j);
  }
  if (cand.empty()) return 0;
  std::stable_sort(cand.begin(), cand.end(), [&](int a, int b) {
    const double da = std::fabs(fs[a] - fs[k]), db = std::fabs(fs[b] - fs[k]);
    return da < db || (da == db && a < b);
  });
  const cpx<double> skk = sp(k);
  if (fs[cand[0]] == fs[k]) {   // the same frequency scale: its solution itself
    const cpx<double> r = cdiv(skk, sp(cand[0]));
    sk[0] = cand[0];
    wk[0] = r.re;
    wk[1] = r.im;
    return 1;
  }
  // distinct scales only (a repeated node would make the weights singular)
  std::vector<int> nodes;
  for (int j : cand) {
    bool dup = false;
    for (int q : nodes) dup |= fs[q] == fs[j];
    if (!dup) nodes.push_back(j);
    if ((int)nodes.size() == nterm) break;
  }
  const int n = (int)nodes.size();
  for (int i = 0; i < n; ++i) {
    double wi = 1.0;
    for (int m = 0; m < n; ++m)
      if (m != i) wi *= (fs[k] - fs[nodes[m]]) / (fs[nodes[i]] - fs[nodes[m]]);
    const cpx<double> r = cdiv(skk, sp(nodes[i]));
    sk[i] = nodes[i];
    wk[2 * i] = wi * r.re;
    wk[2 * i + 1] = wi * r.im;
  }
  return n;
}

// Switches the guesses to the continuation schedule; build_phases derives
// each group's terms from fscale, the spectrum and the chunk plan.
static void build_freq_cont(dogpup_solver* S) {
  S->guess_nterm = S->cont_rb ? kRbTerms : kGuessTerms;
  S->guess_fsrc.clear();
  S->guess_fw.clear();
  S->phases_dirty = true;
  S->clear_graphs();
}

// Per chunk: each group's guess terms as source groups (a term whose
// frequency lies in a later chunk drops the group's guess; one in an earlier
// chunk is read from the caller's x), its level (0
// without terms of other frequencies, else one more than the latest source)
// and the phases: the levels present, each one contiguous range of a copy of
// the work list permuted level by level (a group's items stay together and
// keep their position modulo 8, so their XCD mapping is unchanged).
static int build_phases(dogpup_solver* S) {
  S->chunk_phases.clear();
  S->phases_dirty = false;
  S->guess_self = false;
  const int ng = S->ngroups, nt = S->guess_nterm;
  if (nt <= 0) return DOGPUP_OK;
  std::vector<int> gsrc((size_t)ng * nt, -1), level(ng, 0);
  std::vector<double> gw((size_t)ng * nt * 2, 0.0);
  std::vector<int> pwg, pwt, pwp;
  pwg.reserve(S->nwork);
  pwt.reserve(S->nwork);
  pwp.reserve(S->nwork);
  auto gend = [&](int g) { return g + 1 < ng ? S->h_gibeg[g + 1] : S->nwork; };
  S->chunk_rb.clear();
  for (size_t c = 0; c + 1 < S->chunk_g.size(); ++c) {
    const int g0 = S->chunk_g[c], g1 = S->chunk_g[c + 1];
    bool any_guess = false;
    for (int g = g0; g < g1; ++g) {
      const int m = S->h_gmesh[g], f = S->m_fs0[m] + S->h_gf[g];
      int n = 0;
      bool ok = true;
      if (S->freq_cont) {
        // the schedule: a later-level frequency interpolates the nearest
        // frequencies already solved, of lower levels in this chunk or of an
        // earlier chunk
        const int stride = cont_stride(S);
        const int nfm = S->m_nf[m], k = S->h_gf[g], lk = cont_level(k, nfm, stride);
        auto grp = [&](int j) { return S->m_g0[m] + j * S->m_ncb[m] + S->h_gcb[g]; };
        auto avail = [&](int j) {
          const int sg = grp(j);
          return sg < g0 || (sg < g1 && cont_level(j, nfm, stride) < lk);
        };
        int sk[kMaxGuessTerms];
        double wk[2 * kMaxGuessTerms];
        if (lk > 0)
          n = cont_terms(S->h_fscale.data() + S->m_fs0[m], S->h_spec.empty() ? nullptr : S->h_spec.data() + 2 * S->m_fs0[m],
                         nfm, k, avail, nt, sk, wk);
        for (int q = 0; q < n; ++q) {
          gsrc[(size_t)g * nt + q] = grp(sk[q]);
          gw[2 * ((size_t)g * nt + q)] = wk[2 * q];
          gw[2 * ((size_t)g * nt + q) + 1] = wk[2 * q + 1];
        }
        any_guess |= n > 0;
        continue;
      }
      for (int q = 0; q < nt && ok; ++q) {
        const int sf = S->guess_fsrc[(size_t)f * nt + q];
        if (sf < 0) break;
        const int sg = S->m_g0[m] + (sf - S->m_fs0[m]) * S->m_ncb[m] + S->h_gcb[g];
        if (sg >= g1) ok = false;   // solved after this chunk: no guess (earlier chunks are read from x)
        else {
          gsrc[(size_t)g * nt + n] = sg;
          gw[2 * ((size_t)g * nt + n)] = S->guess_fw[2 * ((size_t)f * nt + q)];
          gw[2 * ((size_t)g * nt + n) + 1] = S->guess_fw[2 * ((size_t)f * nt + q) + 1];
          ++n;
        }
      }
      if (!ok) {
        for (int q = 0; q < nt; ++q) gsrc[(size_t)g * nt + q] = -1;
        n = 0;
      }
      for (int q = 0; q < n; ++q) S->guess_self |= gsrc[(size_t)g * nt + q] == g;
      any_guess |= n > 0;
    }
    // levels: longest chain of other-frequency sources
    for (int round = 0, changed = 1; changed; ++round) {
      DP_ARG(round <= g1 - g0 + 1, "initial-guess terms form a cycle");
      changed = 0;
      for (int g = g0; g < g1; ++g) {
        int L = 0;
        for (int q = 0; q < nt; ++q) {
          const int sg = gsrc[(size_t)g * nt + q];
          if (sg < 0) break;
          if (sg != g && sg >= g0) L = std::max(L, level[sg] + 1);
        }
        if (L != level[g]) {
          level[g] = L;
          changed = 1;
        }
      }
    }
    std::vector<int> lv(level.begin() + g0, level.begin() + g1);
    std::sort(lv.begin(), lv.end());
    lv.erase(std::unique(lv.begin(), lv.end()), lv.end());
    std::vector<dogpup_solver::Phase> ph;
    for (int L : lv) {
      dogpup_solver::Phase p{L, (int)pwg.size(), 0, false, 0};
      for (int g = g0; g < g1; ++g) {
        if (level[g] != L) continue;
        for (int w = S->h_gibeg[g]; w < gend(g); ++w) {
          pwg.push_back(S->h_wgrp[w]);
          pwt.push_back(S->h_wtile[w]);
          pwp.push_back(S->h_wpart[w]);
        }
        p.nw += gend(g) - S->h_gibeg[g];
        p.guess |= gsrc[(size_t)g * nt] >= 0;
        p.nodes += S->h_grows[g];
      }
      ph.push_back(p);
    }
    if (!any_guess) ph.assign(1, dogpup_solver::Phase{0, S->h_gibeg[g0], gend(g1 - 1) - S->h_gibeg[g0], false,
                                                       group_nodes(S, g0, g1)});
    S->chunk_phases.push_back(ph);
    // Galerkin guesses when every term of the chunk is a group of the chunk
    bool rb = S->freq_cont && S->cont_rb && S->mode == DOGPUP_MODE_C64_REFINE && any_guess;
    for (int g = g0; g < g1 && rb; ++g)
      for (int q = 0; q < nt; ++q) rb &= gsrc[(size_t)g * nt + q] < 0 || gsrc[(size_t)g * nt + q] >= g0;
    S->chunk_rb.push_back(rb ? 1 : 0);
  }
  DP_ARG((int)pwg.size() == S->nwork, "phase work list lost items");
  DP_TRY(upload(S->t_pwgrp, pwg.data(), pwg.size() * 4));
  DP_TRY(upload(S->t_pwtile, pwt.data(), pwt.size() * 4));
  DP_TRY(upload(S->t_pwpart, pwp.data(), pwp.size() * 4));
  DP_TRY(upload(S->t_gsrc, gsrc.data(), gsrc.size() * 4));
  DP_TRY(upload(S->t_gw, gw.data(), gw.size() * 8));
  DP_TRY(upload(S->t_glevel, level.data(), level.size() * 4));
  S->h_gsrc = gsrc;
  S->h_pwgrp = pwg;
  DP_TRY(rb_build_plan(S));
  S->clear_graphs();
  return DOGPUP_OK;
}

static int create_batch(dogpup_solver_t* out, int nmesh, const int* h_row0, const int* h_nfreq, const int* h_ncols,
                        int nsp, int mode, int nnzA, const int* d_rptrA, const int* d_cptrA, int nnzG,
                        const int* d_rptrG, const int* d_cptrG, const int* h_order) {
  DP_ARG(out != nullptr, "out is NULL");
  *out = nullptr;
  DP_ARG(nmesh >= 1 && h_row0 && h_nfreq && h_ncols, "bad mesh description");
  DP_ARG(h_row0[0] == 0, "row0[0] must be 0");
  for (int i = 0; i < nmesh; ++i)
    DP_ARG(h_row0[i + 1] > h_row0[i] && h_nfreq[i] > 0 && h_ncols[i] > 0, "mesh %d: bad rows, nfreq or ncols", i);
  const int nn = h_row0[nmesh];
  DP_ARG(nnzA > 0 && nnzG > 0, "bad dimensions");
  DP_ARG(check_nsp(nsp), "group width must be 8, 16, 32 or 64 (got %d)", nsp);
  DP_ARG(mode == DOGPUP_MODE_C128 || mode == DOGPUP_MODE_C64_REFINE || mode == DOGPUP_MODE_C128_PLAIN,
         "unknown mode %d", mode);
  DP_ARG(d_rptrA && d_cptrA && d_rptrG && d_cptrG, "NULL pattern");
  std::unique_ptr<dogpup_solver> S(new dogpup_solver());
  S->nmesh = nmesh; S->nn = nn; S->nsp = nsp; S->mode = mode;
  if (mode == DOGPUP_MODE_C128_PLAIN) S->method = DOGPUP_METHOD_BICGSTAB;
  S->nnzA = nnzA; S->nnzG = nnzG;
  S->rA0 = d_rptrA; S->cA0 = d_cptrA; S->rG0 = d_rptrG; S->cG0 = d_cptrG;
  // Row tile of the kernels (a multiple of 64 rows).  384 in the refinement
  // form, whose continuation levels launch few groups at a time, so that
  // more and shorter work items fill the GPU (ms per step of the 269k-node,
  // 2048-RHS solve + fc2td, same box: 192: 130.5-130.8, 256: 128.9-129.0,
  // 320: 129.1-129.2, 384: 126.6-126.8, 448: 128.1-128.2, 512: 131.8-132.1);
  // 512 in the FP64 forms.  A batch and a single-mesh
  // handle of the same form share the default, so that their solves agree
  // bit for bit; a batch's tile is for the tuner to pick (48 uneven meshes in
  // one default handle: 0.714 s at 256, 0.668-0.691 s at 512).
  const int def_tile = mode == DOGPUP_MODE_C64_REFINE ? 384 : 512;
  S->tile_rows = def_tile;
  if (const char* e = getenv("DOGPUP_TILE")) S->tile_rows = atoi(e);
  if (S->tile_rows < 64 || S->tile_rows % 64) S->tile_rows = def_tile;
  if (const char* e = getenv("DOGPUP_SPMV_LDS_PAD")) S->lds_pad = atoi(e);
  if (const char* e = getenv("DOGPUP_SPMV_UA")) S->unroll_a = atoi(e);
  if (const char* e = getenv("DOGPUP_SPMV_UG")) S->unroll_g = atoi(e);

  // solver node order and the permuted patterns (host-side, once per handle)
  std::vector<int> mesh_of(nn);
  for (int i = 0; i < nmesh; ++i)
    for (int r = h_row0[i]; r < h_row0[i + 1]; ++r) mesh_of[r] = i;
  std::vector<int> order(nn), inv(nn, -1);
  for (int i = 0; i < nn; ++i) order[i] = h_order ? h_order[i] : i;
  for (int i = 0; i < nn; ++i) {
    DP_ARG(order[i] >= 0 && order[i] < nn && inv[order[i]] < 0, "node order is not a permutation");
    DP_ARG(mesh_of[order[i]] == mesh_of[i], "node order moves row %d to another mesh", i);
    inv[order[i]] = i;
  }
  std::vector<int> rA(nn + 1), cA(nnzA), rG(nn + 1), cG(nnzG);
  DP_HIP(hipMemcpy(rA.data(), d_rptrA, (nn + 1) * 4, hipMemcpyDeviceToHost));
  DP_HIP(hipMemcpy(cA.data(), d_cptrA, (size_t)nnzA * 4, hipMemcpyDeviceToHost));
  DP_HIP(hipMemcpy(rG.data(), d_rptrG, (nn + 1) * 4, hipMemcpyDeviceToHost));
  DP_HIP(hipMemcpy(cG.data(), d_cptrG, (size_t)nnzG * 4, hipMemcpyDeviceToHost));
  DP_ARG(rA[0] == 0 && rA[nn] == nnzA && rG[0] == 0 && rG[nn] == nnzG, "CSR row pointers inconsistent with nnz");
  for (int i = 0; i < nn; ++i) DP_ARG(rA[i] <= rA[i + 1] && rG[i] <= rG[i + 1], "CSR row pointers decrease");
  for (int r = 0; r < nn; ++r) {
    for (int p = rA[r]; p < rA[r + 1]; ++p)
      DP_ARG(cA[p] >= 0 && cA[p] < nn && mesh_of[cA[p]] == mesh_of[r], "A column index out of its mesh (row %d)", r);
    for (int p = rG[r]; p < rG[r + 1]; ++p)
      DP_ARG(cG[p] >= 0 && cG[p] < nn && mesh_of[cG[p]] == mesh_of[r], "G column index out of its mesh (row %d)", r);
  }
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
  std::vector<int> rGT(nn + 1, 0), cGT(nnzG), gtmap(nnzG), gtg(nnzG);
  for (int v : cGp) rGT[v + 1]++;
  for (int i = 0; i < nn; ++i) rGT[i + 1] += rGT[i];
  {
    std::vector<int> fill(rGT.begin(), rGT.end() - 1);
    for (int i = 0; i < nn; ++i)
      for (int p = rGp[i]; p < rGp[i + 1]; ++p) {
        const int q = fill[cGp[p]]++;
        cGT[q] = i;
        gtmap[q] = gmap[p];
        gtg[q] = p;
      }
  }
  auto up = [&](dogpup::Buf& b, const std::vector<int>& v) -> int { return upload(b, v.data(), v.size() * 4); };
  // FSAI gather table: every pattern row of at most 8 entries, the A
  // position of each lower-triangle pair found by the same ascending walk of
  // A's row as the searching kernel
  S->fsai_direct = true;
  for (int r = 0; r < nn && S->fsai_direct; ++r) S->fsai_direct = rG[r + 1] - rG[r] <= 8;
  if (S->fsai_direct) {
    std::vector<int> pos((size_t)nn * 36, -1), ginv(nnzG);
    for (int r = 0; r < nn; ++r) {
      const int gs = rG[r], m = rG[r + 1] - gs;
      for (int a = 0; a < m; ++a) {
        const int ra = cG[gs + a], s1 = rA[ra + 1];
        int p = rA[ra];
        for (int b = 0; b <= a; ++b) {
          const int cb = cG[gs + b];
          while (p < s1 && cA[p] < cb) ++p;
          if (p < s1 && cA[p] == cb) pos[(size_t)r * 36 + a * (a + 1) / 2 + b] = p;
        }
      }
    }
    for (int q = 0; q < nnzG; ++q) ginv[gmap[q]] = q;
    DP_TRY(up(S->fsai_pos, pos));
    DP_TRY(up(S->ginv, ginv));
    DP_TRY(up(S->gtg, gtg));
  }
  DP_TRY(up(S->pAr, rAp)); DP_TRY(up(S->pAc, cAp)); DP_TRY(up(S->amap, amap));
  DP_TRY(up(S->pGr, rGp)); DP_TRY(up(S->pGc, cGp)); DP_TRY(up(S->gmap, gmap));
  DP_TRY(up(S->pGTr, rGT)); DP_TRY(up(S->pGTc, cGT)); DP_TRY(up(S->gtmap, gtmap));
  DP_TRY(up(S->order, order));
  S->order_identity = true;
  for (int i = 0; i < nn && S->order_identity; ++i) S->order_identity = order[i] == i;
  S->rA = S->pAr.as<int>(); S->cA = S->pAc.as<int>();
  S->rG = S->pGr.as<int>(); S->cG = S->pGc.as<int>();
  S->rGT = S->pGTr.as<int>(); S->cGT = S->pGTc.as<int>();
  // x-halo lists per 64-row chunk of every mesh
  {
    std::vector<int> chunk_row0;   // first row of every chunk, all meshes
    std::vector<int> mesh_chb(nmesh);
    for (int i = 0; i < nmesh; ++i) {
      mesh_chb[i] = (int)chunk_row0.size();
      for (int r = h_row0[i]; r < h_row0[i + 1]; r += kChunkRows) chunk_row0.push_back(r);
    }
    const int nch = (int)chunk_row0.size();
    auto chunk_end = [&](int c) {
      const int r = chunk_row0[c];
      return std::min(r + kChunkRows, h_row0[mesh_of[r] + 1]);
    };
    const std::vector<int>* pats[3][2] = {{&rAp, &cAp}, {&rGp, &cGp}, {&rGT, &cGT}};
    // per chunk: its distinct columns ascending (the halo), and each
    // nonzero's position in that list; stamp[col] = the last chunk that
    // listed col, slot[col] = its position there
    std::vector<int> stamp(nn), slot(nn);
    for (int m = 0; m < 3; ++m) {
      const std::vector<int>& rp = *pats[m][0];
      const std::vector<int>& cp = *pats[m][1];
      std::vector<int> hp(nch + 1, 0), hc;
      std::vector<unsigned char> hl(cp.size(), 0), mok(nmesh, 1);
      std::vector<int> cols, mcols, mptr;
      std::fill(stamp.begin(), stamp.end(), -1);
      hc.reserve(cp.size() / 3);
      // a mesh fits when every chunk's halo and nonzeros fit the LDS; its
      // chunks are consecutive, so each mesh is listed and then kept or
      // dropped as a whole
      for (int i = 0; i < nmesh; ++i) {
        const int c0 = mesh_chb[i], c1 = i + 1 < nmesh ? mesh_chb[i + 1] : nch;
        mcols.clear();
        mptr.assign(1, 0);
        for (int c = c0; c < c1 && mok[i]; ++c) {
          const int e0 = rp[chunk_row0[c]], e1 = rp[chunk_end(c)];
          cols.clear();
          for (int q = e0; q < e1; ++q)
            if (stamp[cp[q]] != c) {
              stamp[cp[q]] = c;
              cols.push_back(cp[q]);
            }
          if (cols.empty() || (int)cols.size() > kHaloCap || e1 - e0 > kChunkCap) {
            mok[i] = 0;
            break;
          }
          std::sort(cols.begin(), cols.end());
          for (int k = 0; k < (int)cols.size(); ++k) slot[cols[k]] = k;
          for (int q = e0; q < e1; ++q) hl[q] = (unsigned char)slot[cp[q]];
          mcols.insert(mcols.end(), cols.begin(), cols.end());
          mptr.push_back((int)mcols.size());
        }
        if (mok[i]) {
          for (int c = c0; c < c1; ++c) hp[c + 1] = (int)hc.size() + mptr[c - c0 + 1];
          hc.insert(hc.end(), mcols.begin(), mcols.end());
        } else {
          const int e0 = rp[chunk_row0[c0]], e1 = rp[chunk_end(c1 - 1)];
          std::fill(hl.begin() + e0, hl.begin() + e1, (unsigned char)0);
          for (int c = c0; c < c1; ++c) hp[c + 1] = (int)hc.size();
        }
      }
      S->halo_ok[m] = std::find(mok.begin(), mok.end(), 1) != mok.end();
      S->mesh_ok_tmp[m] = mok;
      if (S->halo_ok[m]) {
        if (hc.empty()) hc.push_back(0);
        DP_TRY(up(S->h_ptr[m], hp));
        DP_TRY(up(S->h_col[m], hc));
        DP_TRY(upload(S->h_loc[m], hl.data(), hl.size()));
      }
    }
    S->mesh_chb_tmp = mesh_chb;
  }

  // meshes and groups
  S->m_row0.assign(h_row0, h_row0 + nmesh + 1);
  S->m_nf.assign(h_nfreq, h_nfreq + nmesh);
  S->m_ncols.assign(h_ncols, h_ncols + nmesh);
  long long gbase = 0, vb = 0;
  int fs = 0, rhs = 0, g = 0;
  for (int i = 0; i < nmesh; ++i) {
    const int ncb = (h_ncols[i] + nsp - 1) / nsp, rows = h_row0[i + 1] - h_row0[i];
    const int pG0 = rG[h_row0[i]], nnzGi = rG[h_row0[i + 1]] - pG0;
    S->m_ncb.push_back(ncb);
    S->m_g0.push_back(g);
    S->m_pA0.push_back(rA[h_row0[i]]);
    S->m_pG0.push_back(pG0);
    S->m_nnzG.push_back(nnzGi);
    S->m_gbase.push_back(gbase);
    S->m_fs0.push_back(fs);
    S->m_rhs0.push_back(rhs);
    for (int f = 0; f < h_nfreq[i]; ++f)
      for (int cb = 0; cb < ncb; ++cb, ++g) {
        S->h_grow0.push_back(h_row0[i]);
        S->h_grows.push_back(rows);
        S->h_gvb.push_back(vb);
        S->h_ggv.push_back(gbase + (long long)f * nnzGi - pG0);
        S->h_gcols.push_back(std::min(nsp, h_ncols[i] - cb * nsp));
        S->h_gmesh.push_back(i);
        S->h_gf.push_back(f);
        S->h_gcb.push_back(cb);
        vb += rows;
      }
    gbase += (long long)h_nfreq[i] * nnzGi;
    fs += h_nfreq[i];
    rhs += h_nfreq[i] * h_ncols[i];
  }
  S->ngroups = g;
  S->gvals = gbase;
  S->nf_total = fs;
  S->nrhs_out = rhs;
  DP_ARG(S->ngroups <= (1 << 30), "too many groups (%d)", S->ngroups);
  DP_TRY(upload(S->t_grow0, S->h_grow0.data(), g * 4));
  DP_TRY(upload(S->t_grows, S->h_grows.data(), g * 4));
  DP_TRY(upload(S->t_gvb, S->h_gvb.data(), g * 8));
  DP_TRY(upload(S->t_ggv, S->h_ggv.data(), g * 8));
  DP_TRY(upload(S->t_gcols, S->h_gcols.data(), g * 4));
  {
    std::vector<int> gchb(g);
    for (int q = 0; q < g; ++q) gchb[q] = S->mesh_chb_tmp[S->h_gmesh[q]];
    DP_TRY(upload(S->t_gchb, gchb.data(), g * 4));
    S->mesh_chb_tmp.clear();
    for (int m = 0; m < 3; ++m) {
      std::vector<unsigned char> gok(g);
      for (int q = 0; q < g; ++q) gok[q] = S->mesh_ok_tmp[m][S->h_gmesh[q]];
      DP_TRY(upload(S->t_gok[m], gok.data(), g));
      S->mesh_ok_tmp[m].clear();
    }
  }
  if (const char* e = getenv("DOGPUP_HALO")) S->use_halo = atoi(e) != 0;
  if (const char* e = getenv("DOGPUP_HALO_HALVES")) S->halo_halves = atoi(e) != 0;
  if (const char* e = getenv("DOGPUP_HALO_INTERLEAVE")) S->halo_ilv = atoi(e) != 0;
  if (const char* e = getenv("DOGPUP_CG_RFREE")) S->cg_rfree = atoi(e) != 0;
  S->freq_cont = mode == DOGPUP_MODE_C64_REFINE;
  if (const char* e = getenv("DOGPUP_FREQ_CONT")) S->freq_cont = S->freq_cont && atoi(e) != 0;
  DP_TRY(S->t_gfsc.ensure((size_t)g * 8));
  DP_TRY(S->fscale.ensure((size_t)fs * 8));
  DP_TRY(build_work(S.get()));
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

}  // namespace dogpup

extern "C" {

int dogpup_solver_create(dogpup_solver_t* out, int nn, int nfreq, int ncols, int nsp, int mode,
                         int nnzA, const int* d_rptrA, const int* d_cptrA, int nnzG,
                         const int* d_rptrG, const int* d_cptrG, const int* h_order) {
  DP_ARG(nn > 0 && nfreq > 0 && ncols > 0, "bad dimensions");
  const int row0[2] = {0, nn};
  return create_batch(out, 1, row0, &nfreq, &ncols, nsp, mode, nnzA, d_rptrA, d_cptrA, nnzG, d_rptrG, d_cptrG,
                      h_order);
}

int dogpup_batch_create(dogpup_solver_t* out, int nmesh, const int* h_row0, const int* h_nfreq,
                        const int* h_ncols, int nsp, int mode, int nnzA, const int* d_rptrA, const int* d_cptrA,
                        int nnzG, const int* d_rptrG, const int* d_cptrG, const int* h_order) {
  return create_batch(out, nmesh, h_row0, h_nfreq, h_ncols, nsp, mode, nnzA, d_rptrA, d_cptrA, nnzG, d_rptrG,
                      d_cptrG, h_order);
}

int dogpup_solver_destroy(dogpup_solver_t s) {
  if (s) {
    (void)hipDeviceSynchronize();
    delete s;
  }
  return DOGPUP_OK;
}

// FSAI factors in the reference node order (pattern = the caller's tril(A))
// into g64, FP64, one launch per mesh: mesh i's values for frequency f at
// m_gbase[i] + f*nnzG_i + (p - pG0_i).
static int fsai_reference_order(dogpup_solver* S, hipStream_t st) {
  DP_TRY(S->g64.ensure((size_t)S->gvals * 16));
  for (int i = 0; i < S->nmesh; ++i) {
    FsaiOut fo{S->g64.as<double>() + 2 * (S->m_gbase[i] - S->m_pG0[i]), nullptr, S->m_nnzG[i], FSAI_OUT_SOLVER};
    DP_TRY(fsai_launch(S->a64re.as<double>(), S->a64im.as<double>(), 0, S->rA0, S->cA0, S->rG0 + S->m_row0[i],
                       S->cG0, S->m_row0[i + 1] - S->m_row0[i], S->m_nf[i], S->fscale.as<double>() + S->m_fs0[i],
                       fo, st));
  }
  S->g64_valid = true;
  return DOGPUP_OK;
}

int dogpup_solver_set_matrix(dogpup_solver_t S, const double* d_valA_re, const double* d_valA_im,
                             const double* d_fscale, void* stream) {
  DP_ARG(S && d_valA_re && d_valA_im && d_fscale, "NULL argument");
  hipStream_t st = as_stream(stream);
  // The captured graphs read the values below from these buffers and take
  // their schedule from fscale: they stay valid across a call that keeps
  // both (new optical properties at the same frequencies, as every LM step
  // and every benchmark step makes), so only a change of either clears them.
  // A call that fails leaves the handle without a matrix.
  auto graph_bufs = [S]() {
    return std::vector<void*>{S->a64pre.p, S->a64pim.p, S->a32re.p, S->a32im.p, S->g32.p,
                              S->gt32.p,   S->gp64.p,   S->gtp64.p, S->t_gfsc.p, S->fscale.p};
  };
  const std::vector<void*> bufs0 = graph_bufs();
  const bool had_matrix = S->have_matrix;
  S->have_matrix = false;
  const size_t nA = S->nnzA;
  DP_TRY(S->a64re.ensure(nA * 8));
  DP_TRY(S->a64im.ensure(nA * 8));
  DP_HIP(hipMemcpyAsync(S->a64re.p, d_valA_re, nA * 8, hipMemcpyDeviceToDevice, st));
  DP_HIP(hipMemcpyAsync(S->a64im.p, d_valA_im, nA * 8, hipMemcpyDeviceToDevice, st));
  DP_HIP(hipMemcpyAsync(S->fscale.p, d_fscale, (size_t)S->nf_total * 8, hipMemcpyDeviceToDevice, st));
  // per-group frequency scale
  std::vector<double> hf(S->nf_total), gf(S->ngroups);
  DP_HIP(hipMemcpyAsync(hf.data(), d_fscale, hf.size() * 8, hipMemcpyDeviceToHost, st));
  DP_HIP(hipStreamSynchronize(st));
  for (int g = 0; g < S->ngroups; ++g) gf[g] = hf[S->m_fs0[S->h_gmesh[g]] + S->h_gf[g]];
  DP_HIP(hipMemcpy(S->t_gfsc.p, gf.data(), gf.size() * 8, hipMemcpyHostToDevice));
  const bool same_fscale = had_matrix && hf == S->h_fscale;
  if (!same_fscale) {
    S->clear_graphs();
    S->h_fscale = hf;
    if (S->freq_cont) build_freq_cont(S);
  }
  // FSAI factors: straight into the solver layouts through the gather table
  // (DOGPUP_FSAI_GATHER=1, rows longer than 8 or the pivoted kernel: in the
  // reference order, then gathered)
  const size_t nG = (size_t)S->gvals;
  const char* eg = getenv("DOGPUP_FSAI_GATHER");
  const char* ep = getenv("DOGPUP_FSAI_PIVOT");
  const bool direct = S->fsai_direct && !(eg && atoi(eg)) && !(ep && atoi(ep));
  S->g64_valid = false;
  if (!direct) DP_TRY(fsai_reference_order(S, st));
  // values in solver node order
  DP_TRY(S->a64pre.ensure(nA * 8));
  DP_TRY(S->a64pim.ensure(nA * 8));
  k_gather_real<double><<<grid1(nA), 256, 0, st>>>(S->a64re.as<double>(), S->amap.as<int>(), S->a64pre.as<double>(), nA);
  k_gather_real<double><<<grid1(nA), 256, 0, st>>>(S->a64im.as<double>(), S->amap.as<int>(), S->a64pim.as<double>(), nA);
  DP_LAUNCH_CHECK();
  auto gather_g = [&](auto* tag, Buf& dst, const Buf& map) -> int {
    using TO = std::remove_pointer_t<decltype(tag)>;
    DP_TRY(dst.ensure(nG * 2 * sizeof(TO)));
    if (direct) return DOGPUP_OK;
    for (int i = 0; i < S->nmesh; ++i) {
      const long long n = (long long)S->m_nnzG[i] * S->m_nf[i];
      k_gather_cplx<double, TO><<<grid1(n), 256, 0, st>>>(S->g64.as<double>(), map.as<int>(), dst.as<TO>(), S->m_pG0[i],
                                                  S->m_nnzG[i], S->m_nf[i], S->m_gbase[i]);
      DP_LAUNCH_CHECK();
    }
    return DOGPUP_OK;
  };
  if (S->mode == DOGPUP_MODE_C64_REFINE) {
    DP_TRY(S->a32re.ensure(nA * 4));
    DP_TRY(S->a32im.ensure(nA * 4));
    k_gather_real<float><<<grid1(nA), 256, 0, st>>>(S->a64re.as<double>(), S->amap.as<int>(), S->a32re.as<float>(), nA);
    k_gather_real<float><<<grid1(nA), 256, 0, st>>>(S->a64im.as<double>(), S->amap.as<int>(), S->a32im.as<float>(), nA);
    DP_LAUNCH_CHECK();
    DP_TRY(gather_g((float*)nullptr, S->g32, S->gmap));
    DP_TRY(gather_g((float*)nullptr, S->gt32, S->gtmap));
  } else {
    DP_TRY(gather_g((double*)nullptr, S->gp64, S->gmap));
    DP_TRY(gather_g((double*)nullptr, S->gtp64, S->gtmap));
  }
  if (direct) {
    // G in solver order from the FSAI kernel, then G^T gathered from it
    auto direct_g = [&](auto* tag, Buf& g, Buf& gt) -> int {
      using TO = std::remove_pointer_t<decltype(tag)>;
      for (int i = 0; i < S->nmesh; ++i) {
        FsaiDirect fo{g.as<TO>() + 2 * S->m_gbase[i], S->ginv.as<int>(), S->m_nnzG[i], S->m_pG0[i],
                      std::is_same<TO, float>::value ? 1 : 0};
        DP_TRY(fsai_pos_launch(S->a64re.as<double>(), S->a64im.as<double>(),
                               S->fsai_pos.as<int>() + 36LL * S->m_row0[i], S->rG0 + S->m_row0[i],
                               S->m_row0[i + 1] - S->m_row0[i], S->m_nf[i], S->fscale.as<double>() + S->m_fs0[i],
                               fo, st));
        const long long n = (long long)S->m_nnzG[i] * S->m_nf[i];
        k_gather_cplx<TO, TO><<<grid1(n), 256, 0, st>>>(g.as<TO>(), S->gtg.as<int>(), gt.as<TO>(), S->m_pG0[i],
                                                        S->m_nnzG[i], S->m_nf[i], S->m_gbase[i]);
        DP_LAUNCH_CHECK();
      }
      return DOGPUP_OK;
    };
    if (S->mode == DOGPUP_MODE_C64_REFINE) DP_TRY(direct_g((float*)nullptr, S->g32, S->gt32));
    else DP_TRY(direct_g((double*)nullptr, S->gp64, S->gtp64));
  }
  if (graph_bufs() != bufs0) S->clear_graphs();
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
  // b and x in place of each other (or overlapping) take the permutation
  // buffers, which read b before x is written
  {
    const uintptr_t bytes = (uintptr_t)(S->h_gvb.back() + S->h_grows.back()) * (uintptr_t)S->nsp * 16;
    const uintptr_t xa = reinterpret_cast<uintptr_t>(d_x), ba = reinterpret_cast<uintptr_t>(d_b);
    S->ident_solve = S->order_identity && (xa + bytes <= ba || ba + bytes <= xa);
  }
  const double* xg = S->ident_solve ? d_x : nullptr;
  if (S->mode == DOGPUP_MODE_C128 && xg != S->graph_x) {
    S->clear_graphs();
    S->graph_x = xg;
  }
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
  const int nv_old = nvec_for(S->mode, S->method);
  S->method = method;
  // the refinement form's COCG carves half the vectors: re-plan the chunks
  if (nvec_for(S->mode, method) != nv_old) DP_TRY(build_work(S));
  return DOGPUP_OK;
}

int dogpup_solver_set_rhs_tol(dogpup_solver_t S, const double* h_tol, int n) {
  DP_ARG(S != nullptr, "NULL solver");
  if (h_tol == nullptr || n == 0) {
    S->have_rhs_tol = false;
    return DOGPUP_OK;
  }
  DP_ARG(S->mode != DOGPUP_MODE_C128_PLAIN, "the reference-structure form has one lockstep tolerance");
  DP_ARG(n == S->nrhs_out, "h_tol must hold one tolerance per right-hand side (%d), got %d", S->nrhs_out, n);
  std::vector<double> t((size_t)S->ngroups * S->nsp, 0.0);
  for (int g = 0; g < S->ngroups; ++g)
    for (int c = 0; c < S->nsp; ++c) {
      const int o = out_index(S, g, c);
      if (o < 0) continue;
      DP_ARG(h_tol[o] >= 0.0 && std::isfinite(h_tol[o]), "tolerance %d is %g", o, h_tol[o]);
      t[(size_t)g * S->nsp + c] = h_tol[o];
    }
  DP_TRY(upload(S->rhs_tol, t.data(), t.size() * 8));
  S->have_rhs_tol = true;
  return DOGPUP_OK;
}

int dogpup_solver_set_guess(dogpup_solver_t S, int nterm, const int* h_src, const double* h_w) {
  DP_ARG(S != nullptr, "NULL solver");
  DP_ARG(nterm >= 0 && nterm <= kMaxGuessTerms, "nterm must be 0..%d (got %d)", kMaxGuessTerms, nterm);
  S->freq_cont = false;
  S->phases_dirty = true;
  S->clear_graphs();
  if (nterm == 0) {
    S->guess_nterm = 0;
    S->guess_fsrc.clear();
    S->guess_fw.clear();
    return DOGPUP_OK;
  }
  DP_ARG(S->mode == DOGPUP_MODE_C64_REFINE, "initial guesses are for the refinement form");
  DP_ARG(h_src && h_w, "NULL guess terms");
  std::vector<int> fmesh(S->nf_total);
  for (int i = 0; i < S->nmesh; ++i)
    for (int f = 0; f < S->m_nf[i]; ++f) fmesh[S->m_fs0[i] + f] = i;
  for (int f = 0; f < S->nf_total; ++f)
    for (int q = 0; q < nterm; ++q) {
      const int sf = h_src[(size_t)f * nterm + q];
      if (sf < 0) break;
      DP_ARG(sf < S->nf_total && fmesh[sf] == fmesh[f], "guess term %d of frequency %d names frequency %d of another mesh",
             q, f, sf);
      DP_ARG(std::isfinite(h_w[2 * ((size_t)f * nterm + q)]) && std::isfinite(h_w[2 * ((size_t)f * nterm + q) + 1]),
             "guess weight %d of frequency %d is not finite", q, f);
    }
  S->guess_nterm = nterm;
  S->guess_fsrc.assign(h_src, h_src + (size_t)S->nf_total * nterm);
  S->guess_fw.assign(h_w, h_w + (size_t)S->nf_total * nterm * 2);
  return DOGPUP_OK;
}

int dogpup_solver_set_freq_continuation(dogpup_solver_t S, int on, const double* h_spectrum) {
  DP_ARG(S != nullptr, "NULL solver");
  DP_ARG(on == 0 || on == DOGPUP_CONT_GALERKIN || on == DOGPUP_CONT_LAGRANGE, "unknown continuation kind %d", on);
  DP_ARG(!on || S->mode == DOGPUP_MODE_C64_REFINE, "frequency continuation is for the refinement form");
  if (h_spectrum) {
    for (int f = 0; f < 2 * S->nf_total; ++f) DP_ARG(std::isfinite(h_spectrum[f]), "spectrum entry %d is not finite", f / 2);
    S->h_spec.assign(h_spectrum, h_spectrum + 2 * (size_t)S->nf_total);
  } else {
    S->h_spec.clear();
  }
  S->guess_nterm = 0;
  S->guess_fsrc.clear();
  S->guess_fw.clear();
  S->freq_cont = on != 0;
  if (on) S->cont_rb = on == DOGPUP_CONT_GALERKIN;
  S->phases_dirty = true;
  S->clear_graphs();
  if (S->freq_cont && !S->h_fscale.empty()) build_freq_cont(S);
  return DOGPUP_OK;
}

int dogpup_solver_set_tuning(dogpup_solver_t S, int tile_rows, int unroll_a, int unroll_g, int graph_k) {
  DP_ARG(S != nullptr, "NULL solver");
  DP_ARG(tile_rows == 0 || (tile_rows >= 64 && tile_rows % 64 == 0 && tile_rows <= 65536),
         "tile_rows must be a multiple of 64 (got %d)", tile_rows);
  DP_ARG(unroll_a == 0 || unroll_a == 4 || unroll_a == 8 || unroll_a == 16, "unroll_a must be 4, 8 or 16");
  DP_ARG(unroll_g == 0 || unroll_g == 4 || unroll_g == 8 || unroll_g == 16, "unroll_g must be 4, 8 or 16");
  DP_ARG(graph_k >= 0, "graph_k must be >= 0");
  if (unroll_a) S->unroll_a = unroll_a;
  if (unroll_g) S->unroll_g = unroll_g;
  if (graph_k) S->graph_k = graph_k;
  S->clear_graphs();
  if (tile_rows && tile_rows != S->tile_rows) {
    S->tile_rows = tile_rows;
    DP_TRY(build_work(S));
  }
  return DOGPUP_OK;
}

int dogpup_solver_set_cont_stride(dogpup_solver_t S, int stride) {
  DP_ARG(S != nullptr, "NULL solver");
  DP_ARG(stride == 0 || (stride >= 2 && !(stride & (stride - 1))), "stride must be a power of two >= 2 (got %d)",
         stride);
  if (stride != S->cont_stride) {
    S->cont_stride = stride;
    S->phases_dirty = true;
    S->clear_graphs();
  }
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
  hipStream_t st = as_stream(stream);
  if (!S->g64_valid) DP_TRY(fsai_reference_order(S, st));
  DP_HIP(hipMemcpyAsync(d_out, S->g64.p, (size_t)S->gvals * 16, hipMemcpyDeviceToDevice, st));
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
