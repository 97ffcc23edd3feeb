// Fourier-series transforms, detector sampling and the time-gated adjoint
// Jacobian.
//
// All three transforms are a small dense complex matrix W applied along the
// leading axis of a tall block X (NIN x ncol, columns contiguous):
//   td2fc  (DOGPUP/td2fc.m:26-32): W[k][t] = dt/P * exp(-i f_k t), P = t(end)
//   fc2td  (fc2td.m:24-31):        y = 2 Re(W X), W[t][k] = w_k exp(i f_k t)
//   fc2tg  (fc2tg.m:25-32):        y = Nt 2 Re(W X), W[g][k] = w_k conj(Gate_gk)
// with w_0 = 1/2 (the halved DC term) and w_k = 1 otherwise.  W is built on
// the device; the apply kernel keeps a tile of W rows in LDS and streams X
// once with coalesced column accesses.
//
// Jacobian (dMesh.m:206-247 followed by fc2tg, wghtd_LM_recon.m:58-61):
//   J_TG(m, g, v) = scale(m,g) * sum_k Re( C_gk * phi(v,k,s_m) * phiA(v,k,d_m) )
// with C_gk = -2 Nt w_k conj(Gate_gk) dxyz^3 / s_fpsf(k) supplied by the host.
// One block per node stages that node's forward and adjoint fluence of every
// frequency in LDS (complex64) and every lane produces whole (m, g) rows, so
// the 282 GB complex Jacobian of the reference is never materialised.
#include "common.hpp"

namespace dogpup {

enum { W_TD2FC = 0, W_FC2TD = 1, W_FC2TG = 2 };

__global__ void k_build_w(double* __restrict__ W, int kind, const double* __restrict__ f,
                          const double* __restrict__ t, const double* __restrict__ gates, int nout,
                          int nin, int nt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nout * nin) return;
  const int o = i / nin, k = i % nin;
  double re, im;
  if (kind == W_TD2FC) {
    const double dt = t[1] - t[0], P = t[nt - 1];
    double s, c;
    sincos(-f[o] * t[k], &s, &c);
    re = dt * c / P;
    im = dt * s / P;
  } else if (kind == W_FC2TD) {
    double s, c;
    sincos(f[k] * t[o], &s, &c);
    const double w = (k == 0) ? 0.5 : 1.0;
    re = w * c;
    im = w * s;
  } else {
    const double w = (k == 0) ? 0.5 : 1.0;
    re = w * gates[2 * (o * nin + k)];
    im = -w * gates[2 * (o * nin + k) + 1];
  }
  W[2 * i] = re;
  W[2 * i + 1] = im;
}

constexpr int OT = 8;    // outputs per thread
constexpr int KCH = 256; // W columns staged in LDS per pass

// Y[o][col] = scale * op(sum_i W[o][i] X[i][col]); X real (x_complex = 0) or
// complex; op 0: complex result, op 1: 2*Re (real result).
__global__ __launch_bounds__(256) void k_apply_w(const double* __restrict__ W, const double* __restrict__ X,
                                                 double* __restrict__ Y, int nout, int nin,
                                                 long long ncol, int x_complex, int op,
                                                 double scale) {
  __shared__ double sw[OT * 2 * KCH];
  const int o0 = blockIdx.y * OT;
  const int nrow = min(OT, nout - o0);
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double ar[OT], ai[OT];
#pragma unroll
  for (int q = 0; q < OT; ++q) { ar[q] = 0.0; ai[q] = 0.0; }
  for (int i0 = 0; i0 < nin; i0 += KCH) {
    const int ni = min(KCH, nin - i0);
    __syncthreads();
    for (int j = threadIdx.x; j < OT * ni; j += blockDim.x) {
      const int q = j / ni, i = j % ni;
      if (q < nrow) {
        sw[2 * (q * KCH + i)] = W[2 * ((long long)(o0 + q) * nin + i0 + i)];
        sw[2 * (q * KCH + i) + 1] = W[2 * ((long long)(o0 + q) * nin + i0 + i) + 1];
      }
    }
    __syncthreads();
    if (col < ncol) {
      for (int i = 0; i < ni; ++i) {
        double xr, xi = 0.0;
        if (x_complex) {
          const double2 v = reinterpret_cast<const double2*>(X)[(long long)(i0 + i) * ncol + col];
          xr = v.x;
          xi = v.y;
        } else {
          xr = X[(long long)(i0 + i) * ncol + col];
        }
#pragma unroll
        for (int q = 0; q < OT; ++q) {
          const double wr = sw[2 * (q * KCH + i)], wi = sw[2 * (q * KCH + i) + 1];
          ar[q] += wr * xr - wi * xi;
          ai[q] += wr * xi + wi * xr;
        }
      }
    }
  }
  if (col >= ncol) return;
#pragma unroll
  for (int q = 0; q < OT; ++q) {
    if (q >= nrow) break;
    const long long o = (long long)(o0 + q) * ncol + col;
    if (op == 0) {
      Y[2 * o] = scale * ar[q];
      Y[2 * o + 1] = scale * ai[q];
    } else {
      Y[o] = scale * 2.0 * ar[q];
    }
  }
}

static int apply_w(int kind, const double* X, double* Y, const double* f, const double* t,
                   const double* gates, int nout, int nin, int nt, long long ncol, int x_complex,
                   int op, double scale, hipStream_t st) {
  if (nout == 0 || ncol == 0) return DOGPUP_OK;
  double* W = nullptr;
  DP_HIP(hipMalloc((void**)&W, (size_t)nout * nin * 16));
  k_build_w<<<div_up((long long)nout * nin, 256), 256, 0, st>>>(W, kind, f, t, gates, nout, nin, nt);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) {
    dim3 grid(div_up(ncol, 256), div_up(nout, OT));
    k_apply_w<<<grid, 256, 0, st>>>(W, X, Y, nout, nin, ncol, x_complex, op, scale);
    e = hipGetLastError();
  }
  hipError_t e2 = hipStreamSynchronize(st);
  (void)hipFree(W);
  if (e != hipSuccess) return fail(DOGPUP_ERR_HIP, "transform launch: %s", hipGetErrorString(e));
  if (e2 != hipSuccess) return fail(DOGPUP_ERR_HIP, "transform: %s", hipGetErrorString(e2));
  return DOGPUP_OK;
}

// data(m, k) = sum_q dw[d_m][q] * phi(dnode[d_m][q], k, s_m); phi in solver
// layout; out NM x NF complex, m fastest (MATLAB layout).
__global__ void k_meas(const double* __restrict__ phi, double* __restrict__ out,
                       const int* __restrict__ dnodes, const double* __restrict__ dw,
                       const int* __restrict__ link, int nn, int nf, int ncols, int nsp, int nm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nm * nf) return;
  const int m = i % nm, k = i / nm;
  const int s = link[2 * m], d = link[2 * m + 1];
  const int ncb = (ncols + nsp - 1) / nsp;
  const long long g = (long long)k * ncb + s / nsp;
  const int c = s % nsp;
  double re = 0.0, im = 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double w = dw[4 * d + q];
    if (w == 0.0) continue;
    const long long o = (g * nn + dnodes[4 * d + q]) * nsp + c;
    re += w * phi[2 * o];
    im += w * phi[2 * o + 1];
  }
  out[2 * i] = re;
  out[2 * i + 1] = im;
}

// J_TG[v][g][m] (float).  field: forward and adjoint columns of one solve
// (complex128 solver layout, ncols columns per frequency, group width nsp);
// link[m] = (column of source s_m, column of detector d_m) in that field.
template <int MAXG>
__global__ __launch_bounds__(256) void k_jacobian_tg(
    const double* __restrict__ field, const double* __restrict__ coef, const int* __restrict__ link,
    const float* __restrict__ scale, float* __restrict__ out, int nn, int nf, int ng, int ncols,
    int nsp, int nm) {
  extern __shared__ float2 sm[];  // [nf][ncols] then coefficients [ng][nf]
  float2* sc = sm + nf * ncols;
  const int v = blockIdx.x;
  const int ncb = (ncols + nsp - 1) / nsp;
  for (int j = threadIdx.x; j < nf * ncols; j += blockDim.x) {
    const int k = j / ncols, c = j % ncols;
    const long long o = (((long long)k * ncb + c / nsp) * nn + v) * nsp + c % nsp;
    sm[j] = make_float2((float)field[2 * o], (float)field[2 * o + 1]);
  }
  for (int j = threadIdx.x; j < ng * nf; j += blockDim.x)
    sc[j] = make_float2((float)coef[2 * j], (float)coef[2 * j + 1]);
  __syncthreads();
  for (int m = threadIdx.x; m < nm; m += blockDim.x) {
    const int s = link[2 * m], d = link[2 * m + 1];
    float acc[MAXG];
#pragma unroll
    for (int g = 0; g < MAXG; ++g) acc[g] = 0.f;
    for (int k = 0; k < nf; ++k) {
      const float2 a = sm[k * ncols + s], b = sm[k * ncols + d];
      const float pr = a.x * b.x - a.y * b.y, pi = a.x * b.y + a.y * b.x;
#pragma unroll
      for (int g = 0; g < MAXG; ++g) {
        if (g < ng) {
          const float2 c = sc[g * nf + k];
          acc[g] += c.x * pr - c.y * pi;
        }
      }
    }
    float* o = out + (long long)v * ng * nm;
#pragma unroll
    for (int g = 0; g < MAXG; ++g)
      if (g < ng) o[(long long)g * nm + m] = scale ? acc[g] * scale[(long long)g * nm + m] : acc[g];
  }
}

}  // namespace dogpup

using namespace dogpup;

extern "C" {

int dogpup_td2fc(const double* d_in, double* d_out, const double* d_f, const double* d_t, int nt,
                 int nf, int ncol, int in_complex, void* stream) {
  DP_ARG(d_in && d_out && d_f && d_t && nt >= 2 && nf >= 1 && ncol >= 0, "bad arguments");
  return apply_w(W_TD2FC, d_in, d_out, d_f, d_t, nullptr, nf, nt, nt, ncol, in_complex, 0, 1.0,
                 as_stream(stream));
}

int dogpup_fc2td(const double* d_in, double* d_out, const double* d_f, const double* d_t, int nf,
                 int nt, int ncol, void* stream) {
  DP_ARG(d_in && d_out && d_f && d_t && nt >= 1 && nf >= 1 && ncol >= 0, "bad arguments");
  return apply_w(W_FC2TD, d_in, d_out, d_f, d_t, nullptr, nt, nf, nt, ncol, 1, 1, 1.0,
                 as_stream(stream));
}

int dogpup_fc2tg(const double* d_in, double* d_out, const double* d_gates, int nf, int ng, int ncol,
                 double nt_factor, void* stream) {
  DP_ARG(d_in && d_out && d_gates && nf >= 1 && ng >= 1 && ncol >= 0, "bad arguments");
  return apply_w(W_FC2TG, d_in, d_out, nullptr, nullptr, d_gates, ng, nf, 0, ncol, 1, 1, nt_factor,
                 as_stream(stream));
}

int dogpup_meas(const double* d_phi, double* d_out, const int* d_dnodes, const double* d_dw,
                const int* d_link, int nn, int nf, int ncols, int nsp, int nm, int nd, void* stream) {
  DP_ARG(d_phi && d_out && d_dnodes && d_dw && d_link && nn > 0 && nf > 0 && ncols > 0 && nm >= 0 &&
             nd > 0 && nsp > 0,
         "bad arguments");
  if (nm == 0) return DOGPUP_OK;
  k_meas<<<div_up((long long)nm * nf, 256), 256, 0, as_stream(stream)>>>(d_phi, d_out, d_dnodes, d_dw,
                                                                        d_link, nn, nf, ncols, nsp, nm);
  DP_LAUNCH_CHECK();
  return DOGPUP_OK;
}

int dogpup_jacobian_tg(const double* d_field, const double* d_coef, const int* d_link,
                       const float* d_scale, float* d_out, int nn, int nf, int ng, int ncols, int nsp,
                       int nm, int nv, void* stream) {
  DP_ARG(d_field && d_coef && d_link && d_out, "NULL argument");
  DP_ARG(nn > 0 && nf > 0 && ng > 0 && ncols > 0 && nsp > 0 && nm > 0 && nv > 0 && nv <= nn,
         "bad dimensions");
  DP_ARG(ng <= 16, "jacobian_tg supports at most 16 gates");
  const size_t shm = (size_t)(nf * ncols + ng * nf) * sizeof(float2);
  DP_ARG(shm <= 60 * 1024, "nf*(ncols+ng) too large for one node's LDS stage (%zu bytes)", shm);
  hipStream_t st = as_stream(stream);
  if (ng <= 8)
    k_jacobian_tg<8><<<nv, 256, shm, st>>>(d_field, d_coef, d_link, d_scale, d_out, nn, nf, ng, ncols,
                                           nsp, nm);
  else
    k_jacobian_tg<16><<<nv, 256, shm, st>>>(d_field, d_coef, d_link, d_scale, d_out, nn, nf, ng, ncols,
                                            nsp, nm);
  DP_LAUNCH_CHECK();
  return DOGPUP_OK;
}

}  // extern "C"
