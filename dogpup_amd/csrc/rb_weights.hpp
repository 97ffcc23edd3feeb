// Weights of the Galerkin initial guesses of the frequency continuation
// (host code, no HIP): for one column, the combination of the window's
// converged solutions that solves A_k x = b_k in their span.  Shared by the
// solver (solver.hip) and a CPU test driver.
#pragma once
#include <algorithm>
#include <cmath>
#include <complex>

namespace dogpup {

constexpr int kRbWeightsMax = 8;   // largest window

// Galerkin weights of one column: guess = sum_i y[i] x_i over the n window
// solutions.  G: x_i^H M x_j (n x n row-major, Hermitian), cq: x_i^H q, sj
// and fj: spectrum and scale of each solution, (sk, fk) of the target
// (A_k = K + i fk M, b_k = sk q).  The window is reduced by a pivoted
// Cholesky factorisation G(P, P) = L L^H, pivoting while the largest
// remaining diagonal exceeds 1e-14 of the largest diagonal (the solutions
// are converged to ~1e-7, so directions below ~1e-7 of their norm carry
// noise), so that W = X_P L^-H is M-orthonormal; then (W^H A_k W) z =
// W^H b_k by partial-pivot elimination and y_P = L^-H z.  y = 0 when G
// vanishes or the reduced system is singular.
inline void rb_weights(int n, const std::complex<double>* G, const std::complex<double>* cq,
                       const std::complex<double>* sj, const double* fj, std::complex<double> sk, double fk,
                       std::complex<double>* y) {
  using cd = std::complex<double>;
  constexpr int K = kRbWeightsMax;
  constexpr double eps = 1e-14;
  for (int i = 0; i < n; ++i) y[i] = 0.0;
  double dmax = 0.0;
  for (int i = 0; i < n; ++i) dmax = std::max(dmax, G[i * n + i].real());
  if (!(dmax > 0.0) || !std::isfinite(dmax)) return;
  cd A[K * K], L[K * K];
  int p[K];
  for (int i = 0; i < n * n; ++i) {
    A[i] = G[i];
    L[i] = 0.0;
  }
  for (int i = 0; i < n; ++i) p[i] = i;
  int r = 0;
  for (int k = 0; k < n; ++k) {
    int j = k;
    for (int i = k + 1; i < n; ++i)
      if (A[i * n + i].real() > A[j * n + j].real()) j = i;
    if (!(A[j * n + j].real() > eps * dmax)) break;
    if (j != k) {
      for (int i = 0; i < n; ++i) std::swap(A[k * n + i], A[j * n + i]);
      for (int i = 0; i < n; ++i) std::swap(A[i * n + k], A[i * n + j]);
      for (int i = 0; i < n; ++i) std::swap(L[k * n + i], L[j * n + i]);
      std::swap(p[k], p[j]);
    }
    const double lkk = std::sqrt(A[k * n + k].real());
    L[k * n + k] = lkk;
    for (int i = k + 1; i < n; ++i) L[i * n + k] = A[i * n + k] / lkk;
    for (int i = k + 1; i < n; ++i)
      for (int l = k + 1; l < n; ++l) A[i * n + l] -= L[i * n + k] * std::conj(L[l * n + k]);
    r = k + 1;
  }
  if (r == 0) return;
  // T = L^-H (upper triangular, r x r): L^H T = I by back substitution
  cd T[K * K];
  for (int col = 0; col < r; ++col)
    for (int i = r - 1; i >= 0; --i) {
      cd v = i == col ? cd(1.0) : cd(0.0);
      for (int l = i + 1; l < r; ++l) v -= std::conj(L[l * n + i]) * T[l * r + col];
      T[i * r + col] = v / L[i * n + i].real();
    }
  // the system on the pivots, H_il = x_i^H A_k x_l, then T^H H T z = T^H rhs
  cd H[K * K], HT[K * K], Hr[K * K], rr[K], z[K];
  for (int i = 0; i < r; ++i)
    for (int l = 0; l < r; ++l) {
      const int a = p[i], b = p[l];
      H[i * r + l] = sj[b] * cq[a] + cd(0.0, fk - fj[b]) * G[a * n + b];
    }
  for (int i = 0; i < r; ++i)
    for (int l = 0; l < r; ++l) {
      cd v = 0.0;
      for (int m = 0; m < r; ++m) v += H[i * r + m] * T[m * r + l];
      HT[i * r + l] = v;
    }
  for (int i = 0; i < r; ++i) {
    cd v = 0.0;
    for (int m = 0; m < r; ++m) v += std::conj(T[m * r + i]) * (sk * cq[p[m]]);
    rr[i] = v;
    for (int l = 0; l < r; ++l) {
      cd h = 0.0;
      for (int m = 0; m < r; ++m) h += std::conj(T[m * r + i]) * HT[m * r + l];
      Hr[i * r + l] = h;
    }
  }
  for (int k = 0; k < r; ++k) {
    int j = k;
    for (int i = k + 1; i < r; ++i)
      if (std::abs(Hr[i * r + k]) > std::abs(Hr[j * r + k])) j = i;
    if (!(std::abs(Hr[j * r + k]) > 0.0)) return;
    if (j != k) {
      for (int l = 0; l < r; ++l) std::swap(Hr[k * r + l], Hr[j * r + l]);
      std::swap(rr[k], rr[j]);
    }
    for (int i = k + 1; i < r; ++i) {
      const cd f = Hr[i * r + k] / Hr[k * r + k];
      for (int l = k; l < r; ++l) Hr[i * r + l] -= f * Hr[k * r + l];
      rr[i] -= f * rr[k];
    }
  }
  for (int i = r - 1; i >= 0; --i) {
    cd v = rr[i];
    for (int l = i + 1; l < r; ++l) v -= Hr[i * r + l] * z[l];
    z[i] = v / Hr[i * r + i];
  }
  for (int i = 0; i < r; ++i) {
    cd v = 0.0;
    for (int l = 0; l < r; ++l) v += T[i * r + l] * z[l];
    if (!std::isfinite(v.real()) || !std::isfinite(v.imag())) {
      for (int q = 0; q < n; ++q) y[q] = 0.0;
      return;
    }
    y[p[i]] = v;
  }
}

}  // namespace dogpup
