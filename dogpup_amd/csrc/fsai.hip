// FSAI factor setup on gfx950.
//
// For every (pattern row i, frequency k) the factor row g_i of G_k ~ L_k^-1
// solves A_k[P_i, P_i] x = e_m (P_i ascending, diagonal last, m = |P_i|) and
// is scaled to g_i = x / sqrt(x_m) with the principal complex square root
// (behaviour of DOGPUP/cuda_lib/mex_sources/FSAIP_CUDA.cu:92-160: symmetric
// local matrix gathered from the lower triangle, Gaussian elimination with
// partial pivoting on the modulus, back substitution; A_k = Re A + i*k*Im A,
// FSAIP_CUDA.cu:132-133).
//
// MI355X design: one lane per (row, frequency); the lane's m x (m+1) complex
// augmented system lives in LDS laid out [entry][lane] so the 64 lanes of a
// wavefront hit 64 different banks.  Rows are bucketed by m on the host side
// of the launch (one launch per LDS size class), so a Kuhn mesh (m <= 8)
// needs a single class.
#include "common.hpp"
#include "fsai.hpp"

#include <cstdlib>

namespace dogpup {

template <int MMAX, int BS>
__global__ __launch_bounds__(BS) void k_fsai_rows(
    const double* __restrict__ are, const double* __restrict__ aim, int a_il,
    const int* __restrict__ rptrA, const int* __restrict__ cptrA, const int* __restrict__ rptrG,
    const int* __restrict__ cptrG, int nrows, int nfreqs, const double* __restrict__ fscale,
    int mlo, FsaiOut out) {
  constexpr int W = MMAX + 1;
  __shared__ double sre[MMAX * W * BS];
  __shared__ double sim[MMAX * W * BS];
  const int tid = threadIdx.x;
  const long long t = (long long)blockIdx.x * BS + tid;
  if (t >= (long long)nrows * nfreqs) return;
  const int row = (int)(t / nfreqs);
  const int k = (int)(t % nfreqs);
  const int gs = rptrG[row];
  const int m = rptrG[row + 1] - gs;
  if (m <= mlo || m > MMAX) return;
  const double fk = fscale ? fscale[k] : (double)k;
#define AR(r, c) sre[((r) * W + (c)) * BS + tid]
#define AI(r, c) sim[((r) * W + (c)) * BS + tid]
  for (int r = 0; r < m; ++r)
    for (int c = 0; c <= m; ++c) { AR(r, c) = 0.0; AI(r, c) = 0.0; }
  // gather A_k[P, P] from the lower-triangle entries (P[a] >= P[b] for b <= a)
  for (int a = 0; a < m; ++a) {
    const int ra = cptrG[gs + a];
    const int s0 = rptrA[ra], s1 = rptrA[ra + 1];
    int p = s0;
    for (int b = 0; b <= a; ++b) {
      const int cb = cptrG[gs + b];
      // columns of row ra are ascending and P is ascending: resume the scan
      while (p < s1 && cptrA[p] < cb) ++p;
      if (p < s1 && cptrA[p] == cb) {
        double vr, vi;
        if (a_il) { vr = are[2LL * p]; vi = are[2LL * p + 1]; }
        else { vr = are[p]; vi = aim[p]; }
        vi *= fk;
        AR(a, b) = vr; AI(a, b) = vi;
        AR(b, a) = vr; AI(b, a) = vi;
      }
    }
  }
  AR(m - 1, m) = 1.0;
  // forward elimination with partial pivoting on the modulus
  for (int i = 0; i < m; ++i) {
    int piv = i;
    double best = sqrt(AR(i, i) * AR(i, i) + AI(i, i) * AI(i, i));
    for (int j = i + 1; j < m; ++j) {
      double mj = sqrt(AR(j, i) * AR(j, i) + AI(j, i) * AI(j, i));
      if (mj > best) { best = mj; piv = j; }
    }
    if (piv != i) {
      for (int c = i; c <= m; ++c) {
        double tr = AR(i, c), ti = AI(i, c);
        AR(i, c) = AR(piv, c); AI(i, c) = AI(piv, c);
        AR(piv, c) = tr; AI(piv, c) = ti;
      }
    }
    const cpx<double> d{AR(i, i), AI(i, i)};
    for (int j = i + 1; j < m; ++j) {
      const cpx<double> f = cdiv(cpx<double>{AR(j, i), AI(j, i)}, d);
      for (int c = i; c <= m; ++c) {
        const cpx<double> pr = cmul(f, cpx<double>{AR(i, c), AI(i, c)});
        AR(j, c) -= pr.re;
        AI(j, c) -= pr.im;
      }
    }
  }
  // back substitution; the solution overwrites the augmented column
  for (int i = m - 1; i >= 0; --i) {
    double sr = 0.0, si = 0.0;
    for (int j = i + 1; j < m; ++j) {
      const cpx<double> pr = cmul(cpx<double>{AR(i, j), AI(i, j)}, cpx<double>{AR(j, m), AI(j, m)});
      sr += pr.re;
      si += pr.im;
    }
    const cpx<double> x = cdiv(cpx<double>{AR(i, m) - sr, AI(i, m) - si}, cpx<double>{AR(i, i), AI(i, i)});
    AR(i, m) = x.re;
    AI(i, m) = x.im;
  }
  // principal square root of x_m
  const double xr = AR(m - 1, m), xi = AI(m - 1, m);
  const double mod = sqrt(xr * xr + xi * xi);
  cpx<double> q;
  q.re = sqrt((mod + xr) / 2);
  if (xi == 0.0) {
    if (xr < 0.0) { q.re = 0.0; q.im = sqrt(-xr); }
    else q.im = 0.0;
  } else {
    q.im = (xi > 0 ? 1.0 : -1.0) * sqrt((mod - xr) / 2);
  }
  for (int i = 0; i < m; ++i) {
    const cpx<double> g = cdiv(cpx<double>{AR(i, m), AI(i, m)}, q);
    const long long j = gs + i;
    if (out.layout == FSAI_OUT_REF) {
      out.re[k + (long long)nfreqs * j] = g.re;
      out.im[k + (long long)nfreqs * j] = g.im;
    } else {
      double* o = out.re + 2 * ((long long)k * out.nnzG + j);
      o[0] = g.re;
      o[1] = g.im;
    }
  }
#undef AR
#undef AI
}

// Rows with m <= 8 (every row of a Kuhn mesh): the local system in
// registers.  A_k[P, P] is complex symmetric with a positive definite real
// part (stiffness + Robin + k*i*mass), so its LDL^T factorisation needs no
// pivoting and gives the same solution as the pivoted elimination above to
// roundoff.  The system is padded to 8 x 8 with the identity (which leaves
// the first m unknowns unchanged), every loop is unrolled over compile-time
// indices, and the 36 lower-triangle entries and the unknowns stay in VGPRs:
// 2 waves per SIMD instead of the LDS form's 73.7 KB per 64 lanes (half a
// wave per SIMD).  Only e_m's solution is needed: with L unit lower
// triangular, L y = e_m gives y = e_m, so x_m = 1/d_m and
// x_i = -sum_{j>i} L_ji x_j.
// ld(a, b, vr, vi): entry (a, b), b <= a, of A's lower triangle on P (real
// and imaginary part before the frequency scale); false where A has no
// entry.  st(i, g): factor entry i.
template <typename Ld, typename St>
__device__ inline void fsai_ldl8_row(int m, double fk, Ld&& ld, St&& st) {
  constexpr int M = 8;
  // packed lower triangle: entry (a, b), b <= a, at a(a+1)/2 + b
  double lr[M * (M + 1) / 2], li[M * (M + 1) / 2];
#pragma unroll
  for (int a = 0; a < M; ++a)
#pragma unroll
    for (int b = 0; b <= a; ++b) {
      lr[a * (a + 1) / 2 + b] = (a == b && a >= m) ? 1.0 : 0.0;
      li[a * (a + 1) / 2 + b] = 0.0;
    }
  // gather A_k[P, P] from the lower-triangle entries (P ascending)
#pragma unroll
  for (int a = 0; a < M; ++a) {
    if (a < m) {
#pragma unroll
      for (int b = 0; b <= a; ++b) {
        double vr, vi;
        if (ld(a, b, vr, vi)) {
          lr[a * (a + 1) / 2 + b] = vr;
          li[a * (a + 1) / 2 + b] = vi * fk;
        }
      }
    }
  }
  // in-place LDL^T: diagonal -> d_j, strict lower -> L_ij (unconjugated)
#pragma unroll
  for (int j = 0; j < M; ++j) {
    cpx<double> wj[M];   // w_jk = L_jk d_k
#pragma unroll
    for (int q = 0; q < j; ++q)
      wj[q] = cmul(cpx<double>{lr[j * (j + 1) / 2 + q], li[j * (j + 1) / 2 + q]},
                   cpx<double>{lr[q * (q + 1) / 2 + q], li[q * (q + 1) / 2 + q]});
    cpx<double> d{lr[j * (j + 1) / 2 + j], li[j * (j + 1) / 2 + j]};
#pragma unroll
    for (int q = 0; q < j; ++q) {
      const cpx<double> pr = cmul(cpx<double>{lr[j * (j + 1) / 2 + q], li[j * (j + 1) / 2 + q]}, wj[q]);
      d.re -= pr.re;
      d.im -= pr.im;
    }
    lr[j * (j + 1) / 2 + j] = d.re;
    li[j * (j + 1) / 2 + j] = d.im;
#pragma unroll
    for (int i = j + 1; i < M; ++i) {
      cpx<double> v{lr[i * (i + 1) / 2 + j], li[i * (i + 1) / 2 + j]};
#pragma unroll
      for (int q = 0; q < j; ++q) {
        const cpx<double> pr = cmul(cpx<double>{lr[i * (i + 1) / 2 + q], li[i * (i + 1) / 2 + q]}, wj[q]);
        v.re -= pr.re;
        v.im -= pr.im;
      }
      const cpx<double> l = cdiv(v, d);
      lr[i * (i + 1) / 2 + j] = l.re;
      li[i * (i + 1) / 2 + j] = l.im;
    }
  }
  // x_{m-1} = 1 / d_{m-1}; x_i = -sum_{j=i+1}^{m-1} L_ji x_j
  cpx<double> x[M];
#pragma unroll
  for (int i = 0; i < M; ++i) x[i] = cpx<double>{0.0, 0.0};
#pragma unroll
  for (int i = M - 1; i >= 0; --i) {
    if (i == m - 1) {
      x[i] = cdiv(cpx<double>{1.0, 0.0}, cpx<double>{lr[i * (i + 1) / 2 + i], li[i * (i + 1) / 2 + i]});
    } else if (i < m - 1) {
      cpx<double> sacc{0.0, 0.0};
#pragma unroll
      for (int j = i + 1; j < M; ++j) {
        if (j < m) {
          const cpx<double> pr = cmul(cpx<double>{lr[j * (j + 1) / 2 + i], li[j * (j + 1) / 2 + i]}, x[j]);
          sacc.re += pr.re;
          sacc.im += pr.im;
        }
      }
      x[i] = cpx<double>{-sacc.re, -sacc.im};
    }
  }
  // principal square root of x_m, as the pivoted kernel
  cpx<double> xm{0.0, 0.0};
#pragma unroll
  for (int i = 0; i < M; ++i)
    if (i == m - 1) xm = x[i];
  const double mod = sqrt(xm.re * xm.re + xm.im * xm.im);
  cpx<double> q;
  q.re = sqrt((mod + xm.re) / 2);
  if (xm.im == 0.0) {
    if (xm.re < 0.0) { q.re = 0.0; q.im = sqrt(-xm.re); }
    else q.im = 0.0;
  } else {
    q.im = (xm.im > 0 ? 1.0 : -1.0) * sqrt((mod - xm.re) / 2);
  }
#pragma unroll
  for (int i = 0; i < M; ++i)
    if (i < m) st(i, cdiv(x[i], q));
}

// The register LDL^T form, gathering A_k[P, P] by a walk of A's rows.
template <int BS>
__global__ __launch_bounds__(BS) void k_fsai_ldl8(const double* __restrict__ are, const double* __restrict__ aim,
                                                  int a_il, const int* __restrict__ rptrA,
                                                  const int* __restrict__ cptrA, const int* __restrict__ rptrG,
                                                  const int* __restrict__ cptrG, int nrows, int nfreqs,
                                                  const double* __restrict__ fscale, FsaiOut out) {
  const long long t = (long long)blockIdx.x * BS + threadIdx.x;
  if (t >= (long long)nrows * nfreqs) return;
  const int row = (int)(t / nfreqs);
  const int k = (int)(t % nfreqs);
  const int gs = rptrG[row];
  const int m = rptrG[row + 1] - gs;
  if (m > 8) return;
  const double fk = fscale ? fscale[k] : (double)k;
  // walk row P_a of A once over b = 0..a (columns ascending)
  int wa = -1, p = 0, s1 = 0;
  auto ld = [&](int a, int b, double& vr, double& vi) __attribute__((always_inline)) {
    if (a != wa) {
      wa = a;
      const int ra = cptrG[gs + a];
      p = rptrA[ra];
      s1 = rptrA[ra + 1];
    }
    const int cb = cptrG[gs + b];
    while (p < s1 && cptrA[p] < cb) ++p;
    if (!(p < s1 && cptrA[p] == cb)) return false;
    if (a_il) { vr = are[2LL * p]; vi = are[2LL * p + 1]; }
    else { vr = are[p]; vi = aim[p]; }
    return true;
  };
  auto st = [&](int i, cpx<double> g) __attribute__((always_inline)) {
    const long long j = gs + i;
    if (out.layout == FSAI_OUT_REF) {
      out.re[k + (long long)nfreqs * j] = g.re;
      out.im[k + (long long)nfreqs * j] = g.im;
    } else {
      double* o = out.re + 2 * ((long long)k * out.nnzG + j);
      o[0] = g.re;
      o[1] = g.im;
    }
  };
  fsai_ldl8_row(m, fk, ld, st);
}

// The register LDL^T form with A_k[P, P] gathered through a table of A
// positions (pos[36 * row + a(a+1)/2 + b], -1 where A has no entry), and the
// factor written straight into the solver's G value layout in TO (entry s of
// frequency k at k * nnz + ginv[s] - p0: a row's entries stay together).
// Writing the G^T layout here as well measured 2.5 ms more on 269k rows x 64
// frequencies (each entry goes to another row there); the solver gathers it
// from G instead.
template <int BS, typename TO>
__global__ __launch_bounds__(BS) void k_fsai_ldl8_pos(const double* __restrict__ are,
                                                      const double* __restrict__ aim, const int* __restrict__ pos,
                                                      const int* __restrict__ rptrG, int nrows, int nfreqs,
                                                      const double* __restrict__ fscale, FsaiDirect out) {
  const long long t = (long long)blockIdx.x * BS + threadIdx.x;
  if (t >= (long long)nrows * nfreqs) return;
  const int row = (int)(t / nfreqs);
  const int k = (int)(t % nfreqs);
  const int gs = rptrG[row];
  const int m = rptrG[row + 1] - gs;
  if (m > 8) return;
  const double fk = fscale ? fscale[k] : (double)k;
  const int* pr = pos + 36LL * row;
  auto ld = [&](int a, int b, double& vr, double& vi) __attribute__((always_inline)) {
    const int p = pr[a * (a + 1) / 2 + b];
    if (p < 0) return false;
    vr = are[p];
    vi = aim[p];
    return true;
  };
  TO* og = static_cast<TO*>(out.g);
  auto st = [&](int i, cpx<double> g) __attribute__((always_inline)) {
    const long long jg = (long long)k * out.nnz - out.p0 + out.ginv[gs + i];
    og[2 * jg] = (TO)g.re;
    og[2 * jg + 1] = (TO)g.im;
  };
  fsai_ldl8_row(m, fk, ld, st);
}

int fsai_pos_launch(const double* are, const double* aim, const int* pos, const int* rptrG, int nrows, int nfreqs,
                    const double* fscale, FsaiDirect out, hipStream_t st) {
  if (nrows <= 0 || nfreqs <= 0) return DOGPUP_OK;
  const long long total = (long long)nrows * nfreqs;
  if (out.f32)
    k_fsai_ldl8_pos<128, float><<<div_up(total, 128), 128, 0, st>>>(are, aim, pos, rptrG, nrows, nfreqs, fscale,
                                                                     out);
  else
    k_fsai_ldl8_pos<128, double><<<div_up(total, 128), 128, 0, st>>>(are, aim, pos, rptrG, nrows, nfreqs, fscale,
                                                                      out);
  DP_LAUNCH_CHECK();
  return DOGPUP_OK;
}

__global__ void k_max_row_len(const int* __restrict__ rptr, int n, int* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int m = 0;
  if (i < n) m = rptr[i + 1] - rptr[i];
  // wave reduction then one atomic per wave
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

int fsai_launch(const double* are, const double* aim, int a_il, const int* rptrA,
                const int* cptrA, const int* rptrG, const int* cptrG, int nrows, int nfreqs,
                const double* fscale, FsaiOut out, hipStream_t st) {
  if (nrows <= 0 || nfreqs <= 0) return DOGPUP_OK;
  int* dmax = nullptr;
  DP_HIP(hipMalloc((void**)&dmax, sizeof(int)));
  DP_HIP(hipMemsetAsync(dmax, 0, sizeof(int), st));
  k_max_row_len<<<div_up(nrows, 256), 256, 0, st>>>(rptrG, nrows, dmax);
  DP_LAUNCH_CHECK();
  int hmax = 0;
  DP_HIP(hipMemcpyAsync(&hmax, dmax, sizeof(int), hipMemcpyDeviceToHost, st));
  DP_HIP(hipStreamSynchronize(st));
  DP_HIP(hipFree(dmax));
  if (hmax > 32) return fail(DOGPUP_ERR_UNSUPPORTED, "FSAI pattern row with %d entries (max 32)", hmax);
  const long long total = (long long)nrows * nfreqs;
  static const bool pivoted = [] {
    const char* e = getenv("DOGPUP_FSAI_PIVOT");
    return e && atoi(e) != 0;
  }();
  if (pivoted)
    k_fsai_rows<8, 64><<<div_up(total, 64), 64, 0, st>>>(are, aim, a_il, rptrA, cptrA, rptrG, cptrG,
                                                          nrows, nfreqs, fscale, 0, out);
  else
    k_fsai_ldl8<128><<<div_up(total, 128), 128, 0, st>>>(are, aim, a_il, rptrA, cptrA, rptrG, cptrG,
                                                           nrows, nfreqs, fscale, out);
  DP_LAUNCH_CHECK();
  if (hmax > 8) {
    k_fsai_rows<16, 16><<<div_up(total, 16), 16, 0, st>>>(are, aim, a_il, rptrA, cptrA, rptrG,
                                                          cptrG, nrows, nfreqs, fscale, 8, out);
    DP_LAUNCH_CHECK();
  }
  if (hmax > 16) {
    k_fsai_rows<32, 4><<<div_up(total, 4), 4, 0, st>>>(are, aim, a_il, rptrA, cptrA, rptrG,
                                                       cptrG, nrows, nfreqs, fscale, 16, out);
    DP_LAUNCH_CHECK();
  }
  return DOGPUP_OK;
}

}  // namespace dogpup
