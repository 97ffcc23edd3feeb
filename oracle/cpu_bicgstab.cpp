// Oracle: FSAI setup and FSAI-preconditioned BiCGSTAB for one frequency on
// the host cores (test infrastructure only: the timed CPU baseline and a
// checker; the product never links it).
//
// Restates, in FP64 complex arithmetic with OpenMP over rows:
//  * the FSAI factor rows of FSAIP_CUDA.cu:28-160: for row i with pattern
//    P_i (ascending, diagonal last) the local symmetric matrix A[P_i, P_i]
//    is solved against e_m by partial-pivot Gaussian elimination on the
//    modulus, and g_i = x / sqrt(x_m) with the principal square root
//    (b == 0, a < 0 gives i sqrt(-a));
//  * gpuBicstab_FSAIP, dMesh.m:1283-1341: right preconditioning with
//    M^-1 = G^T G, unconjugated bilinear forms, x0 = 0, r_hat0 = r0, the
//    reciprocal step alph = (v . r_hat0) / rho0, the G-weighted
//    omega = (Gt . Gs) / (Gt . Gt), and lockstep convergence on the largest
//    ||s|| / ||b|| or ||r|| / ||b|| over all columns (tol, maxit).
// Vectors are n x ns row-major (a row's ns complex values together), so a
// sparse row product serves all columns at once.
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include <omp.h>

using cd = std::complex<double>;

namespace {

struct Csr {
  int n;
  const int* rp;
  const int* ci;
  const cd* v;
};

// y = M x over ns columns
void spmm(const Csr& M, const cd* x, cd* y, int ns) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < M.n; ++i) {
    cd* yi = y + (size_t)i * ns;
    for (int c = 0; c < ns; ++c) yi[c] = 0.0;
    for (int p = M.rp[i]; p < M.rp[i + 1]; ++p) {
      const cd a = M.v[p];
      const cd* xj = x + (size_t)M.ci[p] * ns;
      for (int c = 0; c < ns; ++c) yi[c] += a * xj[c];
    }
  }
}

// per-column unconjugated dot sum_i a_ic b_ic
void dots(const cd* a, const cd* b, int n, int ns, cd* out) {
  std::vector<cd> acc((size_t)omp_get_max_threads() * ns, 0.0);
#pragma omp parallel
  {
    cd* mine = acc.data() + (size_t)omp_get_thread_num() * ns;
#pragma omp for schedule(static)
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < ns; ++c) mine[c] += a[(size_t)i * ns + c] * b[(size_t)i * ns + c];
  }
  for (int c = 0; c < ns; ++c) {
    cd s = 0.0;
    for (int t = 0; t < omp_get_max_threads(); ++t) s += acc[(size_t)t * ns + c];
    out[c] = s;
  }
}

void norms(const cd* a, int n, int ns, double* out) {
  std::vector<double> acc((size_t)omp_get_max_threads() * ns, 0.0);
#pragma omp parallel
  {
    double* mine = acc.data() + (size_t)omp_get_thread_num() * ns;
#pragma omp for schedule(static)
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < ns; ++c) mine[c] += std::norm(a[(size_t)i * ns + c]);
  }
  for (int c = 0; c < ns; ++c) {
    double s = 0.0;
    for (int t = 0; t < omp_get_max_threads(); ++t) s += acc[(size_t)t * ns + c];
    out[c] = std::sqrt(s);
  }
}

// principal square root with the b == 0, a < 0 case of FSAIP_CUDA.cu:28-43
cd csqrt_ref(cd z) {
  const double a = z.real(), b = z.imag(), mod = std::sqrt(a * a + b * b);
  const double re = std::sqrt((mod + a) / 2);
  if (b == 0.0) return a < 0.0 ? cd(0.0, std::sqrt(-a)) : cd(re, 0.0);
  return cd(re, (b > 0 ? 1.0 : -1.0) * std::sqrt((mod - a) / 2));
}

// entry (i, j) of the full CSR A (columns ascending), 0 if absent
cd entry(const Csr& A, int i, int j) {
  int lo = A.rp[i], hi = A.rp[i + 1];
  while (lo < hi) {
    const int m = (lo + hi) / 2;
    if (A.ci[m] < j) lo = m + 1;
    else hi = m;
  }
  return (lo < A.rp[i + 1] && A.ci[lo] == j) ? A.v[lo] : cd(0.0);
}

}  // namespace

extern "C" {

// FSAI factor values of A (full symmetric CSR, n rows) on the pattern
// (rG, cG): g[p] for every pattern entry p.  Returns 0, or -1 for a row
// longer than 64 entries.
int oracle_fsai_setup(int n, const int* rA, const int* cA, const double* vA, const int* rG, const int* cG,
                      double* g_out, int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  const Csr A{n, rA, cA, reinterpret_cast<const cd*>(vA)};
  cd* g = reinterpret_cast<cd*>(g_out);
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : bad)
  for (int i = 0; i < n; ++i) {
    const int p0 = rG[i], m = rG[i + 1] - p0;
    if (m > 64 || m < 1) {
      bad = 1;
      continue;
    }
    cd M[64][65];
    for (int a = 0; a < m; ++a) {
      for (int b = 0; b < m; ++b) M[a][b] = entry(A, cG[p0 + a], cG[p0 + b]);
      M[a][m] = (a == m - 1) ? 1.0 : 0.0;
    }
    for (int c = 0; c < m; ++c) {          // partial pivot on the modulus
      int piv = c;
      for (int r = c + 1; r < m; ++r)
        if (std::abs(M[r][c]) > std::abs(M[piv][c])) piv = r;
      if (piv != c)
        for (int k = c; k <= m; ++k) std::swap(M[c][k], M[piv][k]);
      for (int r = c + 1; r < m; ++r) {
        const cd f = M[r][c] / M[c][c];
        for (int k = c; k <= m; ++k) M[r][k] -= f * M[c][k];
      }
    }
    cd x[64];
    for (int r = m - 1; r >= 0; --r) {
      cd s = 0.0;
      for (int k = r + 1; k < m; ++k) s += M[r][k] * x[k];
      x[r] = (M[r][m] - s) / M[r][r];
    }
    const cd q = csqrt_ref(x[m - 1]);
    for (int a = 0; a < m; ++a) g[p0 + a] = x[a] / q;
  }
  return bad ? -1 : 0;
}

// Reference-structure BiCGSTAB on A x = b for ns columns (b, x: n x ns
// row-major complex), G from oracle_fsai_setup on (rG, cG) and its
// transpose (rGT, cGT, vGT).  Outputs the iteration count and the final
// largest relative residual; returns 0.
int oracle_bicgstab(int n, const int* rA, const int* cA, const double* vA, const int* rG, const int* cG,
                    const double* vG, const int* rGT, const int* cGT, const double* vGT, const double* b_in,
                    double* x_out, int ns, double tol, int maxit, int nthreads, int* iters, double* relres) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  const Csr A{n, rA, cA, reinterpret_cast<const cd*>(vA)};
  const Csr G{n, rG, cG, reinterpret_cast<const cd*>(vG)};
  const Csr GT{n, rGT, cGT, reinterpret_cast<const cd*>(vGT)};
  const size_t N = (size_t)n * ns;
  const cd* b = reinterpret_cast<const cd*>(b_in);
  cd* x = reinterpret_cast<cd*>(x_out);
  std::vector<cd> r(b, b + N), rhat(b, b + N), p(b, b + N), h(N), v(N), s(N), u(N), z(N), t(N), w(N), hx(N);
  std::vector<double> err0(ns), e(ns);
  std::vector<cd> rho0(ns), alph(ns), om(ns), rho(ns), beta(ns), d1(ns), d2(ns);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < N; ++q) x[q] = 0.0;
  norms(r.data(), n, ns, err0.data());
  dots(rhat.data(), r.data(), n, ns, rho0.data());
  auto maxrel = [&](const std::vector<cd>& a) {
    norms(a.data(), n, ns, e.data());
    double mx = 0.0;
    for (int c = 0; c < ns; ++c)
      if (err0[c] > 0.0) mx = std::max(mx, e[c] / err0[c]);
    return mx;
  };
  int it = 0;
  double err = 1.0;
  for (it = 1; it <= maxit; ++it) {
    spmm(G, p.data(), u.data(), ns);                    // h = G^T G p
    spmm(GT, u.data(), h.data(), ns);
    spmm(A, h.data(), v.data(), ns);                    // v = A h
    dots(v.data(), rhat.data(), n, ns, d1.data());
    for (int c = 0; c < ns; ++c) alph[c] = d1[c] / rho0[c];
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < N; ++q) {
      const int c = (int)(q % ns);
      hx[q] = x[q] + h[q] / alph[c];
      s[q] = r[q] - v[q] / alph[c];
    }
    err = maxrel(s);
    if (err < tol) {
      std::memcpy(x, hx.data(), N * sizeof(cd));
      break;
    }
    spmm(G, s.data(), u.data(), ns);                    // Gs
    spmm(GT, u.data(), z.data(), ns);                   // z = G^T G s
    spmm(A, z.data(), t.data(), ns);                    // t = A z
    spmm(G, t.data(), w.data(), ns);                    // Gt
    dots(w.data(), u.data(), n, ns, d1.data());
    dots(w.data(), w.data(), n, ns, d2.data());
    for (int c = 0; c < ns; ++c) om[c] = d1[c] / d2[c];
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < N; ++q) {
      const int c = (int)(q % ns);
      x[q] = hx[q] + om[c] * z[q];
      r[q] = s[q] - om[c] * t[q];
      p[q] = p[q] - om[c] * v[q];
    }
    err = maxrel(r);
    if (err < tol) break;
    dots(rhat.data(), r.data(), n, ns, rho.data());
    for (int c = 0; c < ns; ++c) {
      beta[c] = rho[c] / rho0[c] / (alph[c] * om[c]);
      rho0[c] = rho[c];
    }
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < N; ++q) p[q] = r[q] + beta[(int)(q % ns)] * p[q];
  }
  *iters = std::min(it, maxit);
  *relres = err;
  return 0;
}

}  // extern "C"
