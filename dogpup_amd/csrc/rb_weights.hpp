// Weights of the Galerkin initial guesses of the frequency continuation: for
// one column, the combination of the window's converged solutions that
// solves A_k x = b_k in their span.  Plain C++ on interleaved (re, im)
// doubles, so that the solver runs it in a kernel (one thread per target
// column) and a CPU test compiles the same code with g++.
#pragma once
#include <cmath>
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DOGPUP_HD __host__ __device__
#else
#define DOGPUP_HD
#endif

namespace dogpup {

constexpr int kRbWeightsMax = 8;   // largest window

struct RbC {
  double re, im;
};
DOGPUP_HD inline RbC rb_c(double re, double im = 0.0) { return RbC{re, im}; }
DOGPUP_HD inline RbC operator+(RbC a, RbC b) { return RbC{a.re + b.re, a.im + b.im}; }
DOGPUP_HD inline RbC operator-(RbC a, RbC b) { return RbC{a.re - b.re, a.im - b.im}; }
DOGPUP_HD inline RbC operator*(RbC a, RbC b) { return RbC{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
DOGPUP_HD inline RbC operator*(double s, RbC a) { return RbC{s * a.re, s * a.im}; }
DOGPUP_HD inline RbC rb_conj(RbC a) { return RbC{a.re, -a.im}; }
DOGPUP_HD inline double rb_abs2(RbC a) { return a.re * a.re + a.im * a.im; }
DOGPUP_HD inline RbC rb_div(RbC a, RbC b) {
  // scaled by the larger component of b (Smith's method)
  if (std::fabs(b.re) >= std::fabs(b.im)) {
    const double r = b.im / b.re, d = b.re + b.im * r;
    return RbC{(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  }
  const double r = b.re / b.im, d = b.im + b.re * r;
  return RbC{(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}

// Galerkin weights of one column: guess = sum_i y[i] x_i over the n window
// solutions (complex values interleaved re, im).  G: x_i^H M x_j (n x n
// row-major, Hermitian), cq: x_i^H q, sj and fj: spectrum and scale of each
// solution, (sk, fk) of the target (A_k = K + i fk M, b_k = sk q), so that
// x_i^H A_k x_j = s_j x_i^H q + i (fk - f_j) x_i^H M x_j.  The window is
// reduced by a pivoted Cholesky factorisation G(P, P) = L L^H, pivoting
// while the largest remaining diagonal exceeds 1e-14 of the largest
// diagonal (the solutions are converged to ~1e-7, so directions below
// ~1e-7 of their norm carry noise), so that W = X_P L^-H is M-orthonormal;
// then (W^H A_k W) z = W^H b_k by partial-pivot elimination and
// y_P = L^-H z.  y = 0 when G vanishes or the reduced system is singular.
// K: the largest window (a call with n > K returns zeros).
template <int K = kRbWeightsMax>
DOGPUP_HD inline void rb_weights(int n, const double* Gd, const double* cqd, const double* sjd, const double* fj,
                                 double skr, double ski, double fk, double* yd) {
  constexpr double eps = 1e-14;
  const RbC* G = reinterpret_cast<const RbC*>(Gd);
  const RbC* cq = reinterpret_cast<const RbC*>(cqd);
  const RbC* sj = reinterpret_cast<const RbC*>(sjd);
  RbC* y = reinterpret_cast<RbC*>(yd);
  const RbC sk = rb_c(skr, ski);
  for (int i = 0; i < n; ++i) y[i] = rb_c(0.0);
  if (n > K) return;
  double dmax = 0.0;
  for (int i = 0; i < n; ++i) dmax = G[i * n + i].re > dmax ? G[i * n + i].re : dmax;
  if (!(dmax > 0.0) || !std::isfinite(dmax)) return;
  RbC A[K * K], L[K * K];
  int p[K];
  for (int i = 0; i < n * n; ++i) {
    A[i] = G[i];
    L[i] = rb_c(0.0);
  }
  for (int i = 0; i < n; ++i) p[i] = i;
  int r = 0;
  for (int k = 0; k < n; ++k) {
    int j = k;
    for (int i = k + 1; i < n; ++i)
      if (A[i * n + i].re > A[j * n + j].re) j = i;
    if (!(A[j * n + j].re > eps * dmax)) break;
    if (j != k) {
      for (int i = 0; i < n; ++i) {
        const RbC t = A[k * n + i];
        A[k * n + i] = A[j * n + i];
        A[j * n + i] = t;
      }
      for (int i = 0; i < n; ++i) {
        const RbC t = A[i * n + k];
        A[i * n + k] = A[i * n + j];
        A[i * n + j] = t;
      }
      for (int i = 0; i < n; ++i) {
        const RbC t = L[k * n + i];
        L[k * n + i] = L[j * n + i];
        L[j * n + i] = t;
      }
      const int t = p[k];
      p[k] = p[j];
      p[j] = t;
    }
    const double lkk = std::sqrt(A[k * n + k].re);
    L[k * n + k] = rb_c(lkk);
    for (int i = k + 1; i < n; ++i) L[i * n + k] = (1.0 / lkk) * A[i * n + k];
    for (int i = k + 1; i < n; ++i)
      for (int l = k + 1; l < n; ++l) A[i * n + l] = A[i * n + l] - L[i * n + k] * rb_conj(L[l * n + k]);
    r = k + 1;
  }
  if (r == 0) return;
  // T = L^-H (upper triangular, r x r): L^H T = I by back substitution
  RbC T[K * K];
  for (int col = 0; col < r; ++col)
    for (int i = r - 1; i >= 0; --i) {
      RbC v = rb_c(i == col ? 1.0 : 0.0);
      for (int l = i + 1; l < r; ++l) v = v - rb_conj(L[l * n + i]) * T[l * r + col];
      T[i * r + col] = (1.0 / L[i * n + i].re) * v;
    }
  // the system on the pivots, H_il = x_i^H A_k x_l, then T^H H T z = T^H rhs
  RbC H[K * K], HT[K * K], Hr[K * K], rr[K], z[K];
  for (int i = 0; i < r; ++i)
    for (int l = 0; l < r; ++l) {
      const int a = p[i], b = p[l];
      H[i * r + l] = sj[b] * cq[a] + rb_c(0.0, fk - fj[b]) * G[a * n + b];
    }
  for (int i = 0; i < r; ++i)
    for (int l = 0; l < r; ++l) {
      RbC v = rb_c(0.0);
      for (int m = 0; m < r; ++m) v = v + H[i * r + m] * T[m * r + l];
      HT[i * r + l] = v;
    }
  for (int i = 0; i < r; ++i) {
    RbC v = rb_c(0.0);
    for (int m = 0; m < r; ++m) v = v + rb_conj(T[m * r + i]) * (sk * cq[p[m]]);
    rr[i] = v;
    for (int l = 0; l < r; ++l) {
      RbC h = rb_c(0.0);
      for (int m = 0; m < r; ++m) h = h + rb_conj(T[m * r + i]) * HT[m * r + l];
      Hr[i * r + l] = h;
    }
  }
  for (int k = 0; k < r; ++k) {
    int j = k;
    for (int i = k + 1; i < r; ++i)
      if (rb_abs2(Hr[i * r + k]) > rb_abs2(Hr[j * r + k])) j = i;
    if (!(rb_abs2(Hr[j * r + k]) > 0.0)) return;
    if (j != k) {
      for (int l = 0; l < r; ++l) {
        const RbC t = Hr[k * r + l];
        Hr[k * r + l] = Hr[j * r + l];
        Hr[j * r + l] = t;
      }
      const RbC t = rr[k];
      rr[k] = rr[j];
      rr[j] = t;
    }
    for (int i = k + 1; i < r; ++i) {
      const RbC f = rb_div(Hr[i * r + k], Hr[k * r + k]);
      for (int l = k; l < r; ++l) Hr[i * r + l] = Hr[i * r + l] - f * Hr[k * r + l];
      rr[i] = rr[i] - f * rr[k];
    }
  }
  for (int i = r - 1; i >= 0; --i) {
    RbC v = rr[i];
    for (int l = i + 1; l < r; ++l) v = v - Hr[i * r + l] * z[l];
    z[i] = rb_div(v, Hr[i * r + i]);
  }
  RbC out[K];
  for (int i = 0; i < r; ++i) {
    RbC v = rb_c(0.0);
    for (int l = 0; l < r; ++l) v = v + T[i * r + l] * z[l];
    if (!std::isfinite(v.re) || !std::isfinite(v.im)) return;
    out[i] = v;
  }
  for (int i = 0; i < r; ++i) y[p[i]] = out[i];
}

}  // namespace dogpup
