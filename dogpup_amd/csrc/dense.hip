// Dense algebra of the LM step (reconstruction/tik_inv.m, reshape_J.m):
// normal matrices on MFMA (FP32 with FP32 accumulation, or exact FP64
// products of FP32 operands with FP64 accumulation), and an FP64 Cholesky
// factorisation and triangular solves for the damped system.
//
// GEMM tiles: a 256-thread block (4 wavefronts) computes a 64 x 64 tile of
// C; each wavefront owns a 32 x 32 quarter as 2 x 2 tiles of
// v_mfma_f32_16x16x4_f32 (or v_mfma_f64_16x16x4_f64).  Operand tiles of
// 64 x 16 per K step alternate between two LDS buffers, the next step's
// loads in registers while the current one multiplies.  MFMA 16x16x4 lane layout (both
// precisions): lane l supplies A[l%16][l/16] and B[l/16][l%16] (K index
// l/16); its accumulator register r holds D[4*(l/16)+r][l%16] for FP32 and
// D[l/16+4*r][l%16] for FP64 (measured on gfx950 by an exact-integer test).
#include <algorithm>
#include <type_traits>

#include "common.hpp"

namespace dogpup {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int GT = 64;   // C tile
constexpr int GK = 16;   // K step

// C[m][n] = alpha * sum_k A[m*lda + k] * B[n*ldb + k]   (TRANS_B = true, "NT")
// C[m][n] = alpha * sum_k A[m*lda + k] * B[k*ldb + n]   (TRANS_B = false, "NN")
// Inputs FP32.  F64: operands converted exactly to FP64 in LDS and
// multiplied on v_mfma_f64_16x16x4f64 with FP64 accumulation (for damped
// normal equations whose shift is below FP32 rounding); otherwise
// v_mfma_f32_16x16x4f32 with FP32 accumulation.  Output FP64 (OUT_D) or FP32.
typedef double f64x4 __attribute__((ext_vector_type(4)));

// Split K (P != nullptr): block z of grid.z sums k in [z*ks, min((z+1)*ks, K))
// and stores its unscaled partial tile, FP64, at P[(z*M + m)*N + n];
// k_gemm_reduce then adds the slices in the order z = 0, 1, ...
template <bool TRANS_B, bool OUT_D, bool F64>
__global__ __launch_bounds__(256) void k_gemm_f32(const float* __restrict__ A, const float* __restrict__ B,
                                                  void* __restrict__ Cv, int M, int N, int K, long long lda,
                                                  long long ldb, long long ldc, float alpha,
                                                  int lower_only, double* __restrict__ P, int ks) {
  using E = typename std::conditional<F64, double, float>::type;
  using V4 = typename std::conditional<F64, f64x4, f32x4>::type;
  // two LDS buffers: step s is stored into buffer s % 2, so one barrier per
  // step orders the stores after every wave's reads of the step before
  __shared__ E As[2][GT][GK + 1];
  __shared__ E Bs[2][GT][GK + 1];
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  if (lower_only && n0 > m0 + GT - 1) return;  // tile entirely above the diagonal
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  V4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = V4{0, 0, 0, 0};
  const int kb = P ? blockIdx.z * ks : 0;
  const int ke = P ? min(K, kb + ks) : K;
  // Software pipeline: the K step after the current one is loaded into
  // registers (ra, rb) before the current step's MFMAs, so its load latency
  // overlaps them; out-of-range elements are zeros.  Thread t covers A row
  // t/4 (and B row t/4, or B k-row t/16 when not transposed), 4 elements.
  const int ar = t >> 2, akq = (t & 3) * 4;
  const int bkk = t >> 4, bnq = (t & 15) * 4;
  float ra[4], rb[4];
  auto gload = [&](int k0) __attribute__((always_inline)) {
    const int gm = m0 + ar;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int gk = k0 + akq + q;
      ra[q] = (gm < M && gk < ke) ? A[(long long)gm * lda + gk] : 0.f;
    }
    if (TRANS_B) {
      const int gn = n0 + ar;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int gk = k0 + akq + q;
        rb[q] = (gn < N && gk < ke) ? B[(long long)gn * ldb + gk] : 0.f;
      }
    } else {
      const int gk = k0 + bkk;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int gn = n0 + bnq + q;
        rb[q] = (gn < N && gk < ke) ? B[(long long)gk * ldb + gn] : 0.f;
      }
    }
  };
  if (kb < ke) gload(kb);
  int buf = 0;
  for (int k0 = kb; k0 < ke; k0 += GK, buf ^= 1) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      As[buf][ar][akq + q] = (E)ra[q];
      if (TRANS_B) Bs[buf][ar][akq + q] = (E)rb[q];
      else Bs[buf][bnq + q][bkk] = (E)rb[q];
    }
    __syncthreads();
    if (k0 + GK < ke) gload(k0 + GK);
#pragma unroll
    for (int kk = 0; kk < GK; kk += 4) {
      const int ka = kk + (lane >> 4);
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const E a = As[buf][wm + 16 * i + (lane & 15)][ka];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const E b = Bs[buf][wn + 16 * j + (lane & 15)][ka];
          if constexpr (F64) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i][j], 0, 0, 0);
          else acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[i][j], 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // accumulator row: FP32 MFMA 4*(lane/16) + r, FP64 MFMA lane/16 + 4*r
        const int rr = F64 ? ((lane >> 4) + 4 * r) : (4 * (lane >> 4) + r);
        const int gm = m0 + wm + 16 * i + rr;
        const int gn = n0 + wn + 16 * j + (lane & 15);
        if (gm < M && gn < N && (!lower_only || gn <= gm)) {
          if (P) {
            P[((long long)blockIdx.z * M + gm) * N + gn] = (double)acc[i][j][r];
            continue;
          }
          const E v = (E)alpha * acc[i][j][r];
          if (OUT_D) reinterpret_cast<double*>(Cv)[(long long)gm * ldc + gn] = (double)v;
          else reinterpret_cast<float*>(Cv)[(long long)gm * ldc + gn] = (float)v;
        }
      }
}

// C[m][n] = alpha * sum_z P[z][m][n], z in order (the split-K slices).
template <bool OUT_D>
__global__ void k_gemm_reduce(const double* __restrict__ P, int S, int M, int N, void* __restrict__ Cv,
                              long long ldc, double alpha, int lower_only) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)M * N) return;
  const int m = (int)(i / N), n = (int)(i % N);
  if (lower_only && n > m) return;
  double v = 0.0;
  for (int z = 0; z < S; ++z) v += P[(long long)z * M * N + i];
  if (OUT_D) reinterpret_cast<double*>(Cv)[(long long)m * ldc + n] = alpha * v;
  else reinterpret_cast<float*>(Cv)[(long long)m * ldc + n] = (float)(alpha * v);
}

// out[j] = sum_i A[i][j]^2 (column sums of squares of a rows x cols FP32
// matrix, FP64 accumulation) -- one thread per column.
__global__ void k_col_sumsq(const float* __restrict__ A, int rows, long long cols, double* __restrict__ out) {
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  double s = 0.0;
  for (int i = 0; i < rows; ++i) {
    const double v = A[(long long)i * cols + j];
    s += v * v;
  }
  out[j] = s;
}

// out[j] = sum_i A[i][j] * y[i]  (A^T y, FP64 accumulation)
__global__ void k_gemv_t(const float* __restrict__ A, const double* __restrict__ y, int rows, long long cols,
                         double* __restrict__ out) {
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  double s = 0.0;
  for (int i = 0; i < rows; ++i) s += (double)A[(long long)i * cols + j] * y[i];
  out[j] = s;
}

// ---------------------------------------------------------------- Cholesky
// In-place lower Cholesky A = L L^T of an n x n row-major FP64 matrix
// (only the lower triangle is read and written), right-looking by NB-wide
// panels.
constexpr int NB = 32;

// Factor the NB x NB diagonal block at (k0, k0) in LDS; info[0] = first
// non-positive pivot (1-based) or untouched.
__global__ __launch_bounds__(256) void k_potrf_diag(double* __restrict__ A, int n, int k0, int* info) {
  __shared__ double L[NB][NB + 1];
  const int nb = min(NB, n - k0);
  for (int i = threadIdx.x; i < nb * nb; i += blockDim.x) {
    const int r = i / nb, c = i % nb;
    L[r][c] = A[(long long)(k0 + r) * n + k0 + c];
  }
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    if (threadIdx.x == 0) {
      const double d = L[j][j];
      if (!(d > 0.0)) {
        if (info[0] == 0) info[0] = k0 + j + 1;
        L[j][j] = 1.0;
      } else {
        L[j][j] = sqrt(d);
      }
    }
    __syncthreads();
    const double djj = L[j][j];
    for (int r = j + 1 + threadIdx.x; r < nb; r += blockDim.x) L[r][j] /= djj;
    __syncthreads();
    // trailing update inside the block
    for (int i = threadIdx.x; i < (nb - j - 1) * (nb - j - 1); i += blockDim.x) {
      const int r = j + 1 + i / (nb - j - 1), c = j + 1 + i % (nb - j - 1);
      if (c <= r) L[r][c] -= L[r][j] * L[c][j];
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < nb * nb; i += blockDim.x) {
    const int r = i / nb, c = i % nb;
    if (c <= r) A[(long long)(k0 + r) * n + k0 + c] = L[r][c];
  }
}

// Panel below the diagonal block: X = A[k0+nb:, k0:k0+nb] * L_kk^-T, one
// thread per row.
__global__ __launch_bounds__(256) void k_trsm_panel(double* __restrict__ A, int n, int k0) {
  __shared__ double L[NB][NB + 1];
  const int nb = min(NB, n - k0);
  for (int i = threadIdx.x; i < nb * nb; i += blockDim.x) {
    const int r = i / nb, c = i % nb;
    L[r][c] = (c <= r) ? A[(long long)(k0 + r) * n + k0 + c] : 0.0;
  }
  __syncthreads();
  const int row = k0 + nb + blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  double x[NB];
  double* a = A + (long long)row * n + k0;
  for (int j = 0; j < nb; ++j) {
    double s = a[j];
    for (int l = 0; l < j; ++l) s -= x[l] * L[j][l];
    x[j] = s / L[j][j];
  }
  for (int j = 0; j < nb; ++j) a[j] = x[j];
}

// Trailing update A[i][j] -= sum_l P[i][l] P[j][l] for k0+nb <= j <= i,
// P = A[:, k0:k0+nb]; 64 x 64 tiles, FP64.
__global__ __launch_bounds__(256) void k_syrk_update(double* __restrict__ A, int n, int k0) {
  const int nb = min(NB, n - k0);
  const int s0 = k0 + nb;
  const int i0 = s0 + blockIdx.y * 64, j0 = s0 + blockIdx.x * 64;
  if (j0 > i0 + 63 || i0 >= n || j0 >= n) return;
  __shared__ double Pi[64][NB + 1], Pj[64][NB + 1];
  for (int q = threadIdx.x; q < 64 * nb; q += blockDim.x) {
    const int r = q / nb, c = q % nb;
    Pi[r][c] = (i0 + r < n) ? A[(long long)(i0 + r) * n + k0 + c] : 0.0;
    Pj[r][c] = (j0 + r < n) ? A[(long long)(j0 + r) * n + k0 + c] : 0.0;
  }
  __syncthreads();
  // each thread: 4 x 4 outputs
  const int tr = (threadIdx.x / 16) * 4, tc = (threadIdx.x % 16) * 4;
  double acc[4][4] = {};
  for (int l = 0; l < nb; ++l) {
    double a[4], b[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) { a[q] = Pi[tr + q][l]; b[q] = Pj[tc + q][l]; }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[p][q] += a[p] * b[q];
  }
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = i0 + tr + p, j = j0 + tc + q;
      if (i < n && j < n && j <= i) A[(long long)i * n + j] -= acc[p][q];
    }
}

// Triangular solves with nrhs right-hand sides stored n x nrhs row-major.
// Forward: L Y = B.  Block row k: B[k] -= L[k, :k0] Y[:k0] (k_trsv_update),
// then the NB x NB triangle (k_trsv_diag).  Backward uses L^T.
__global__ void k_trsv_update(const double* __restrict__ L, double* __restrict__ B, int n, int nrhs, int k0,
                              int trans) {
  const int nb = min(NB, n - k0);
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)nb * nrhs) return;
  const int r = (int)(idx / nrhs), c = (int)(idx % nrhs);
  const int row = k0 + r;
  // eight partial sums, so that eight products' loads are in flight at once
  // (one running sum waited for each load in turn)
  constexpr int U = 8;
  double ps[U];
#pragma unroll
  for (int u = 0; u < U; ++u) ps[u] = 0.0;
  const int l0 = trans ? k0 + nb : 0, l1 = trans ? n : k0;
  int l = l0;
  for (; l + U <= l1; l += U) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const double lv = trans ? L[(long long)(l + u) * n + row] : L[(long long)row * n + l + u];
      ps[u] += lv * B[(long long)(l + u) * nrhs + c];
    }
  }
  for (; l < l1; ++l) {
    const double lv = trans ? L[(long long)l * n + row] : L[(long long)row * n + l];
    ps[0] += lv * B[(long long)l * nrhs + c];
  }
  double s = 0.0;
#pragma unroll
  for (int u = 0; u < U; ++u) s += ps[u];
  B[(long long)row * nrhs + c] -= s;
}

__global__ __launch_bounds__(256) void k_trsv_diag(const double* __restrict__ L, double* __restrict__ B, int n,
                                                   int nrhs, int k0,
                            int trans) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nrhs) return;
  const int nb = min(NB, n - k0);
  // the thread's column of the block in registers (NB is a compile-time
  // bound and every loop is unrolled), with the same operations in the same
  // order as the in-memory form
  double y[NB];
#pragma unroll
  for (int r = 0; r < NB; ++r) y[r] = r < nb ? B[(long long)(k0 + r) * nrhs + c] : 0.0;
  if (!trans) {
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      if (r < nb) {
        const int row = k0 + r;
        double s = y[r];
#pragma unroll
        for (int l = 0; l < r; ++l) s -= L[(long long)row * n + k0 + l] * y[l];
        y[r] = s / L[(long long)row * n + row];
      }
    }
  } else {
#pragma unroll
    for (int r = NB - 1; r >= 0; --r) {
      if (r < nb) {
        const int row = k0 + r;
        double s = y[r];
#pragma unroll
        for (int l = r + 1; l < NB; ++l)
          if (l < nb) s -= L[(long long)(k0 + l) * n + row] * y[l];
        y[r] = s / L[(long long)row * n + row];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < NB; ++r)
    if (r < nb) B[(long long)(k0 + r) * nrhs + c] = y[r];
}

// A[i][i] += shift
__global__ void k_add_diag(double* __restrict__ A, int n, double shift) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[(long long)i * n + i] += shift;
}

}  // namespace dogpup

using namespace dogpup;

extern "C" {

int dogpup_gemm_f32(const float* d_A, const float* d_B, void* d_C, int M, int N, int K, long long lda,
                    long long ldb, long long ldc, float alpha, int trans_b, int out_f64, int lower_only,
                    int compute_f64, void* stream) {
  DP_ARG(d_A && d_B && d_C && M >= 0 && N >= 0 && K >= 0, "bad arguments");
  if (M == 0 || N == 0) return DOGPUP_OK;
  dim3 grid(div_up(N, GT), div_up(M, GT));
  DP_ARG(grid.y <= 65535, "M too large");
  hipStream_t st = as_stream(stream);
  // Split K when the tiles alone cannot fill the GPU (normal matrices of a
  // few thousand rows over a quarter-million columns): slices of at least
  // kSplitMinK, enough of them for about kSplitBlocks blocks.  A block walks
  // K in 16-wide steps and has only the next step's loads in flight, so the
  // blocks in flight per CU hide the rest of the load latency.
  constexpr long long kSplitBlocks = 4096, kSplitMinK = 2048;
  const long long tiles = lower_only ? (long long)grid.y * (grid.y + 1) / 2 : (long long)grid.x * grid.y;
  int S = (int)std::min<long long>((kSplitBlocks + tiles - 1) / tiles, std::max<long long>(1, K / kSplitMinK));
  S = std::max(1, std::min(S, 64));
  double* P = nullptr;
  int ks = K;
  if (S > 1) {
    ks = (int)(((long long)K + S - 1) / S);
    ks = (ks + GK - 1) / GK * GK;
    S = (int)(((long long)K + ks - 1) / ks);
    // every entry the reduction reads (n <= m when lower_only) is written by
    // each slice's tile, so the buffer needs no clearing
    DP_HIP(hipMallocAsync((void**)&P, (size_t)S * M * N * sizeof(double), st));
    grid.z = S;
  }
#define DP_GEMM(TB, OD, F6)                                                                                  \
  k_gemm_f32<TB, OD, F6><<<grid, 256, 0, st>>>(d_A, d_B, d_C, M, N, K, lda, ldb, ldc, alpha, lower_only, P, \
                                               ks)
  if (compute_f64) {
    if (trans_b) { if (out_f64) DP_GEMM(true, true, true); else DP_GEMM(true, false, true); }
    else { if (out_f64) DP_GEMM(false, true, true); else DP_GEMM(false, false, true); }
  } else {
    if (trans_b) { if (out_f64) DP_GEMM(true, true, false); else DP_GEMM(true, false, false); }
    else { if (out_f64) DP_GEMM(false, true, false); else DP_GEMM(false, false, false); }
  }
#undef DP_GEMM
  DP_LAUNCH_CHECK();
  if (P) {
    const long long n = (long long)M * N;
    if (out_f64) k_gemm_reduce<true><<<div_up(n, 256), 256, 0, st>>>(P, S, M, N, d_C, ldc, alpha, lower_only);
    else k_gemm_reduce<false><<<div_up(n, 256), 256, 0, st>>>(P, S, M, N, d_C, ldc, alpha, lower_only);
    DP_LAUNCH_CHECK();
    DP_HIP(hipFreeAsync(P, st));
  }
  return DOGPUP_OK;
}

int dogpup_col_sumsq(const float* d_A, int rows, long long cols, double* d_out, void* stream) {
  DP_ARG(d_A && d_out && rows >= 0 && cols >= 0, "bad arguments");
  if (cols == 0) return DOGPUP_OK;
  k_col_sumsq<<<div_up(cols, 256), 256, 0, as_stream(stream)>>>(d_A, rows, cols, d_out);
  DP_LAUNCH_CHECK();
  return DOGPUP_OK;
}

int dogpup_gemv_t(const float* d_A, const double* d_y, int rows, long long cols, double* d_out,
                  void* stream) {
  DP_ARG(d_A && d_y && d_out && rows >= 0 && cols >= 0, "bad arguments");
  if (cols == 0) return DOGPUP_OK;
  k_gemv_t<<<div_up(cols, 256), 256, 0, as_stream(stream)>>>(d_A, d_y, rows, cols, d_out);
  DP_LAUNCH_CHECK();
  return DOGPUP_OK;
}

int dogpup_chol_factor(double* d_A, int n, double shift, int* h_info, void* stream) {
  DP_ARG(d_A && n > 0, "bad arguments");
  hipStream_t st = as_stream(stream);
  int* dinfo = nullptr;
  DP_HIP(hipMalloc((void**)&dinfo, sizeof(int)));
  hipError_t e = hipMemsetAsync(dinfo, 0, sizeof(int), st);
  if (e == hipSuccess && shift != 0.0) k_add_diag<<<div_up(n, 256), 256, 0, st>>>(d_A, n, shift);
  for (int k0 = 0; k0 < n && e == hipSuccess; k0 += NB) {
    k_potrf_diag<<<1, 256, 0, st>>>(d_A, n, k0, dinfo);
    const int rest = n - k0 - NB;
    if (rest > 0) {
      k_trsm_panel<<<div_up(rest, 256), 256, 0, st>>>(d_A, n, k0);
      const unsigned nt = div_up(rest, 64);
      k_syrk_update<<<dim3(nt, nt), 256, 0, st>>>(d_A, n, k0);
    }
    e = hipGetLastError();
  }
  int info = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(dinfo);
  if (e != hipSuccess) return fail(DOGPUP_ERR_HIP, "cholesky: %s", hipGetErrorString(e));
  if (h_info) *h_info = info;
  if (info) return fail(DOGPUP_ERR_ARG, "matrix not positive definite (pivot %d)", info);
  return DOGPUP_OK;
}

int dogpup_chol_solve(const double* d_L, double* d_B, int n, int nrhs, void* stream) {
  DP_ARG(d_L && d_B && n > 0 && nrhs > 0, "bad arguments");
  hipStream_t st = as_stream(stream);
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int nb = min(NB, n - k0);
    if (k0 > 0) k_trsv_update<<<div_up((long long)nb * nrhs, 256), 256, 0, st>>>(d_L, d_B, n, nrhs, k0, 0);
    k_trsv_diag<<<div_up(nrhs, 256), 256, 0, st>>>(d_L, d_B, n, nrhs, k0, 0);
  }
  const int last = ((n - 1) / NB) * NB;
  for (int k0 = last; k0 >= 0; k0 -= NB) {
    const int nb = min(NB, n - k0);
    if (k0 + nb < n) k_trsv_update<<<div_up((long long)nb * nrhs, 256), 256, 0, st>>>(d_L, d_B, n, nrhs, k0, 1);
    k_trsv_diag<<<div_up(nrhs, 256), 256, 0, st>>>(d_L, d_B, n, nrhs, k0, 1);
  }
  DP_LAUNCH_CHECK();
  return DOGPUP_OK;
}

}  // extern "C"
