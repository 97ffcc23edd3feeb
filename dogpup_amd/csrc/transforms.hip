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
// The 282 GB complex Jacobian of the reference is never materialised.
//
// fc2td of a whole solver-layout field (k_fc2td_mfma64): every column e of
// the block [k][e] (e = (group, node, column) in the solver order) is a real
// (NT x 2NF) x (2NF x 1) product, y[t][e] = sum_kk Wd[t][kk] X[kk][e] with
// Wd[t][(k,0)] = 2 w_k cos(f_k t), Wd[t][(k,1)] = -2 w_k sin(f_k t) and
// X[(k,0)][e] = Re F_k(e), X[(k,1)][e] = Im F_k(e): one GEMM over all columns
// on v_mfma_f64_16x16x4f64, FP64 products and accumulation, FP32 output.  A
// persistent 512-thread block per CU holds Wd (128 time bins per pass up to
// 64 frequencies, 64 beyond) in LDS; a wavefront keeps all of a pass's bins
// of 16 columns in accumulator registers, so every field value it loads
// feeds 8 (4) MFMAs.
//
// MFMA form (k_jtg_mfma, up to 32 distinct source and 32 distinct detector
// columns): per node and gate the link block is a real GEMM
//   J_g[d][s] = sum_kk A_g[d][kk] B[kk][s],  kk = (k, re/im),
//   A_g[d][(k,0)] = Re(C_gk psi_dk), A_g[d][(k,1)] = -Im(C_gk psi_dk),
//   B[(k,0)][s] = Re phi_sk,         B[(k,1)][s] = Im phi_sk,
// i.e. (32 x 2NF) x (2NF x 32) per gate, on v_mfma_f64_16x16x4f64 (FP64
// operands and accumulation).  Two wavefronts own one node (16 detector rows each): they
// stream the node's fluence two frequencies (one K step of 4) at a time
// straight into the MFMA operand registers and keep all gates' accumulator
// tiles in registers.  Links
// are then picked out of the 32x32 block through an (d, s) -> m table.
// Scalar form (k_jacobian_tg, any link list): one block per node stages that
// node's fluence of every frequency in LDS (complex128) and every lane
// produces whole (m, g) rows, summed in FP64.
#include <algorithm>
#include <mutex>
#include <vector>

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
  extern __shared__ double2 sm[];  // [nf][ncols] then coefficients [ng][nf]
  double2* sc = sm + nf * ncols;
  const int v = blockIdx.x;
  const int ncb = (ncols + nsp - 1) / nsp;
  for (int j = threadIdx.x; j < nf * ncols; j += blockDim.x) {
    const int k = j / ncols, c = j % ncols;
    const long long o = (((long long)k * ncb + c / nsp) * nn + v) * nsp + c % nsp;
    sm[j] = make_double2(field[2 * o], field[2 * o + 1]);
  }
  for (int j = threadIdx.x; j < ng * nf; j += blockDim.x) sc[j] = make_double2(coef[2 * j], coef[2 * j + 1]);
  __syncthreads();
  for (int m = threadIdx.x; m < nm; m += blockDim.x) {
    const int s = link[2 * m], d = link[2 * m + 1];
    double acc[MAXG];
#pragma unroll
    for (int g = 0; g < MAXG; ++g) acc[g] = 0.0;
    for (int k = 0; k < nf; ++k) {
      const double2 a = sm[k * ncols + s], b = sm[k * ncols + d];
      const double pr = a.x * b.x - a.y * b.y, pi = a.x * b.y + a.y * b.x;
#pragma unroll
      for (int g = 0; g < MAXG; ++g) {
        if (g < ng) {
          const double2 c = sc[g * nf + k];
          acc[g] += c.x * pr - c.y * pi;
        }
      }
    }
    float* o = out + (long long)v * ng * nm;
#pragma unroll
    for (int g = 0; g < MAXG; ++g)
      if (g < ng) o[(long long)g * nm + m] = (float)(scale ? acc[g] * (double)scale[(long long)g * nm + m] : acc[g]);
  }
}

// The MFMA kernels load operands unconditionally from clamped, in-bounds
// indices: entries past the frequency range meet zero rows of the DFT/gate
// coefficients, and padding rows/columns are never stored, so the finite
// values they read contribute nothing.  (Selecting zero after the load would
// make the compiler wait for every load at once and defeat the prefetch.)
//
// Their K loops are unrolled so that each prefetch register is reloaded right
// after its own use and read again one unrolled iteration later: values
// rotated through register copies would make the copy wait for the load
// within the same step.  The values entering the loop pass through
// `primed`: if every value reaching a loop-carried register were a plain
// load, the optimiser would turn the register into a rotated address and
// issue the load where the value is used.  (x + 0.0 is not simplified away,
// since it maps -0.0 to +0.0, which changes no product.)  A lane that needs
// only the real or only the imaginary part loads that double alone: picking
// the component of a loaded pair by lane would put the pair in scratch memory.
__device__ inline double2 primed(double2 v) { return make_double2(v.x + 0.0, v.y + 0.0); }
__device__ inline double primed(double v) { return v + 0.0; }

// Wd (double, ntp x kks, zero padded) for time bins t[t0 .. t0 + ntp) of
// fc2td: Wd[r][2k] = 2 w_k cos(f_k t), Wd[r][2k+1] = -2 w_k sin(f_k t).
__global__ void k_build_wd(double* __restrict__ Wd, const double* __restrict__ f, const double* __restrict__ t,
                           int t0, int nt, int nf, int ntp, int kks) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntp * kks) return;
  const int r = i / kks, kk = i % kks, k = kk >> 1;
  double v = 0.0;
  if (t0 + r < nt && k < nf) {
    double sn, cs;
    sincos(f[k] * t[t0 + r], &sn, &cs);
    const double w = (k == 0) ? 1.0 : 2.0;   // 2 w_k with the halved DC term
    v = (kk & 1) ? -w * sn : w * cs;
  }
  Wd[i] = v;
}

// y[t0 + t][e] = sum_kk Wd[t][kk] X[kk][e] for t < min(NTT * 16, nt - t0).
// F: complex128 [nf][E]; out: float [nt][E].  A wavefront computes NTT time
// tiles (16 bins each, the whole pass) of a 16-column tile on
// v_mfma_f64_16x16x4f64 (FP64 products and sums; only the stored value is
// rounded to FP32), with Wd for the pass's NTT*16 bins in LDS.  FP64 because
// a time bin of a curve far from its source is a near-total cancellation of
// the 2nf terms: in FP32 such curves came out wrong by up to their own peak.
// NTT = 8 on 512-thread blocks (2 waves per SIMD, 184 VGPRs) measured 9.8 ms
// on 269,001 x 32 columns x 64 frequencies to 256 bins, against 11.0 ms for
// NTT = 4 with two wavefronts per column tile on 1024-thread blocks: each
// field value loaded feeds twice the MFMAs.
constexpr int kFc2tdThreads = 512;
template <int NTT>
__global__ __launch_bounds__(kFc2tdThreads) void k_fc2td_mfma64(const double* __restrict__ F,
                                                                const double* __restrict__ Wd,
                                                                float* __restrict__ out, long long E, int nf,
                                                                int kks, int t0, int nt) {
  extern __shared__ double sWd[];   // NTT*16 rows x kks
  for (int i = threadIdx.x; i < NTT * 16 * kks; i += kFc2tdThreads) sWd[i] = Wd[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, slot = threadIdx.x >> 6;
  constexpr int slots = kFc2tdThreads / 64;
  const int i16 = lane & 15, kk = lane >> 4, kf = kk >> 1, im = kk & 1;
  const int nfp8 = (nf + 7) & ~7;   // Wd holds zero columns up to 2*nfp8
  const long long ntiles = (E + 15) / 16;
  const double* wbase = sWd + i16 * kks + kk;
  typedef double d4 __attribute__((ext_vector_type(4)));
  for (long long mt = (long long)blockIdx.x * slots + slot; mt < ntiles; mt += (long long)gridDim.x * slots) {
    const long long e = mt * 16 + i16;
    const bool oke = e < E;
    d4 acc[NTT];
#pragma unroll
    for (int i = 0; i < NTT; ++i) acc[i] = d4{0.0, 0.0, 0.0, 0.0};
    // four K steps (eight frequencies) per iteration; each X register is
    // reloaded for the next iteration right after its use
    const long long ec = oke ? e : E - 1;
    const double* Xe = F + 2 * ec + im;      // this lane's component of column ec
    double xq[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) xq[j] = primed(Xe[2 * (long long)min(2 * j + kf, nf - 1) * E]);
    for (int k0 = 0; k0 < nfp8; k0 += 8) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double b = xq[j];
        xq[j] = Xe[2 * (long long)min(k0 + 8 + 2 * j + kf, nf - 1) * E];
        __builtin_amdgcn_sched_barrier(0);   // issue the reload before this step's MFMAs
        const double* wrow = wbase + 2 * (k0 + 2 * j);
#pragma unroll
        for (int i = 0; i < NTT; ++i)
          acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(wrow[i * 16 * kks], b, acc[i], 0, 0, 0);
      }
    }
    // accumulator register r of lane l holds D[l/16 + 4*r][l%16] (the f64
    // 16x16x4 layout)
    if (oke) {
#pragma unroll
      for (int i = 0; i < NTT; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int tt = t0 + i * 16 + kk + 4 * r;
          if (tt < nt) out[(long long)tt * E + e] = (float)acc[i][r];
        }
    }
  }
}

// Shape of the fc2td passes for nf frequencies: ntt time tiles per pass (8
// while the pass's Wd fits the LDS, else 4) and the LDS row stride kks
// (doubles; 2 mod 32, so the 16 rows a wavefront reads start in different
// banks).
struct Fc2tdShape {
  int ntt, kks;
  int bins() const { return ntt * 16; }
  size_t shm() const { return (size_t)bins() * kks * sizeof(double); }
};
static Fc2tdShape fc2td_shape(int nf) {
  const int nfp8 = (nf + 7) & ~7;
  Fc2tdShape s;
  s.kks = ((2 * nfp8 + 31) / 32) * 32 + 2;
  s.ntt = 8;
  if (s.shm() > 160 * 1024) s.ntt = 4;
  return s;
}

// The fc2td passes over all bins: Wd is rebuilt in d_wd (at least
// bins * kks doubles) for each pass; ordered on `st`, asynchronous.
static hipError_t fc2td_field_passes(const double* d_field, float* d_out, const double* d_f, const double* d_t,
                                     int nf, int nt, long long ncol, double* d_wd, int ncu, hipStream_t st) {
  const Fc2tdShape sh = fc2td_shape(nf);
  const int ntp = sh.bins();
  const long long ntiles = (ncol + 15) / 16, slots = kFc2tdThreads / 64;
  const unsigned grid = (unsigned)std::max<long long>(1, std::min<long long>(ncu, (ntiles + slots - 1) / slots));
  const void* fn = sh.ntt == 8 ? (const void*)k_fc2td_mfma64<8> : (const void*)k_fc2td_mfma64<4>;
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh.shm());
  for (int t0 = 0; t0 < nt && e == hipSuccess; t0 += ntp) {
    k_build_wd<<<div_up((long long)ntp * sh.kks, 256), 256, 0, st>>>(d_wd, d_f, d_t, t0, nt, nf, ntp, sh.kks);
    e = hipGetLastError();
    if (e != hipSuccess) break;
    if (sh.ntt == 8)
      k_fc2td_mfma64<8><<<grid, kFc2tdThreads, sh.shm(), st>>>(d_field, d_wd, d_out, ncol, nf, sh.kks, t0, nt);
    else
      k_fc2td_mfma64<4><<<grid, kFc2tdThreads, sh.shm(), st>>>(d_field, d_wd, d_out, ncol, nf, sh.kks, t0, nt);
    e = hipGetLastError();
  }
  return e;
}

// MFMA J-TG.  tab = [dcol(32) | scol(32) | mlink(32x32)]: field columns of
// the rows (detectors) and columns (sources) of the link block (-1 =
// padding) and mlink[i*32 + j] = m of the link (source scol[j], detector
// dcol[i]) or -1; coef: complex128 NG x NF.  Gates [gb, gb + NGB) in this
// launch.  Two wavefronts per node (detector rows 0-15 and 16-31), so a
// wavefront keeps NGB x 2 accumulator tiles; 4 wavefronts per block.
// FP64 throughout (v_mfma_f64_16x16x4f64): for early gates and distant
// pairs J is a near-total cancellation of the frequency terms, which FP32
// products and sums leave at up to 1e-2 of such a column's size (measured on
// the 269k-node, 64-frequency case); only the stored J is rounded to FP32.
template <int NGB>
__global__ __launch_bounds__(256) void k_jtg_mfma(const double* __restrict__ field,
                                                  const double* __restrict__ coef,
                                                  const int* __restrict__ tab, const float* __restrict__ scale,
                                                  float* __restrict__ out, int nn, int nf, int ng, int gb,
                                                  int ncols, int nsp, int nm, int nv) {
  __shared__ double2 sc[NGB * 128];   // C_gk (zero for k >= nf), row stride 128
  __shared__ int sd[32], ss[32];
  __shared__ int sm[32 * 32];
  const int nfp4 = (nf + 3) & ~3;     // K loop: four frequencies per iteration
  for (int i = threadIdx.x; i < NGB * 128; i += 256) {
    const int g = i / 128, k = i % 128;
    double2 c = make_double2(0.0, 0.0);
    if (gb + g < ng && k < nf) c = make_double2(coef[2 * ((gb + g) * nf + k)], coef[2 * ((gb + g) * nf + k) + 1]);
    sc[i] = c;
  }
  if (threadIdx.x < 32) {
    sd[threadIdx.x] = tab[threadIdx.x];
    ss[threadIdx.x] = tab[32 + threadIdx.x];
  }
  for (int i = threadIdx.x; i < 1024; i += 256) sm[i] = tab[64 + i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int v = blockIdx.x * 2 + (wv >> 1), t = wv & 1;
  if (v >= nv) return;
  const int i16 = lane & 15, kk = lane >> 4, kf = kk >> 1, im = kk & 1;
  const int ncb = (ncols + nsp - 1) / nsp;
  // operand columns of this lane: detector t*16 + i16 (A row), sources
  // u*16 + i16 (B columns)
  // padding rows/columns (-1) read column 0; they are never stored
  const int cd = max(sd[t * 16 + i16], 0);
  const long long offd = ((long long)(cd / nsp) * nn + v) * nsp + cd % nsp;
  long long offs[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int cs = max(ss[u * 16 + i16], 0);
    offs[u] = ((long long)(cs / nsp) * nn + v) * nsp + cs % nsp;
  }
  const long long kstride = (long long)ncb * nn * nsp;   // complex elements per frequency
  const double2* F = reinterpret_cast<const double2*>(field);
  typedef double d4 __attribute__((ext_vector_type(4)));
  d4 acc[NGB][2];
#pragma unroll
  for (int g = 0; g < NGB; ++g)
#pragma unroll
    for (int u = 0; u < 2; ++u) acc[g][u] = d4{0.0, 0.0, 0.0, 0.0};
  // two K steps (four frequencies) per iteration, each operand register
  // reloaded for the next iteration right after its use; frequencies past
  // nf read frequency nf-1 and meet C = 0
  // B operands: this lane's component (re or im) of the source fluence
  const double* Fs0 = field + 2 * offs[0] + im;
  const double* Fs1 = field + 2 * offs[1] + im;
  double2 pd[2];
  double ps0[2], ps1[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const long long ko = (long long)min(2 * j + kf, nf - 1) * kstride;
    pd[j] = primed(F[ko + offd]);
    ps0[j] = primed(Fs0[2 * ko]);
    ps1[j] = primed(Fs1[2 * ko]);
  }
  for (int k0 = 0; k0 < nfp4; k0 += 4) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int k = k0 + 2 * j + kf;
      const double b0 = ps0[j], b1 = ps1[j];
      const double dr = pd[j].x, di = pd[j].y;
      const long long ko = (long long)min(k + 4, nf - 1) * kstride;
      pd[j] = F[ko + offd];
      ps0[j] = Fs0[2 * ko];
      ps1[j] = Fs1[2 * ko];
      __builtin_amdgcn_sched_barrier(0);   // issue the reloads before this step's MFMAs
#pragma unroll
      for (int g = 0; g < NGB; ++g) {
        const double2 c = sc[g * 128 + k];
        const double a = im ? -(c.x * di + c.y * dr) : (c.x * dr - c.y * di);
        acc[g][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b0, acc[g][0], 0, 0, 0);
        acc[g][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b1, acc[g][1], 0, 0, 0);
      }
    }
  }
  // accumulator register r of lane l holds D[l/16 + 4*r][l%16] of a tile
  // (the f64 16x16x4 layout; the f32 one holds D[4*(l/16) + r])
#pragma unroll
  for (int g = 0; g < NGB; ++g) {
    if (gb + g < ng) {
      float* o = out + ((long long)v * ng + gb + g) * nm;
      const float* scg = scale ? scale + (long long)(gb + g) * nm : nullptr;
#pragma unroll
      for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = sm[(t * 16 + kk + 4 * r) * 32 + u * 16 + i16];
          if (m >= 0) o[m] = (float)(scg ? acc[g][u][r] * (double)scg[m] : acc[g][u][r]);
        }
    }
  }
}

// Scratch of dogpup_fc2td_field_noalloc: the largest DFT table of one pass
// (128 bins x 130 doubles for nf <= 64, 64 x 258 for nf <= 128), 133,120 B.
constexpr int kFc2tdScratch = 128 * 130;
static __device__ __align__(256) double scratch[kFc2tdScratch];

constexpr int kMaxDevices = 64;
static std::mutex fc2td_scr_mu;
static double* fc2td_scr[kMaxDevices] = {};   // device address of `scratch` per device
static int fc2td_ncu[kMaxDevices] = {};

// The current device's copy of `scratch` (resolved on the first call on the
// device, with that device current) and its CU count.
static int fc2td_scratch(double** out, int* ncu) {
  int dev = 0;
  DP_HIP(hipGetDevice(&dev));
  DP_ARG(dev >= 0 && dev < kMaxDevices, "device %d out of range", dev);
  std::lock_guard<std::mutex> lk(fc2td_scr_mu);
  if (!fc2td_scr[dev]) {
    void* p = nullptr;
    DP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    int n = 0;
    DP_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    DP_ARG(p != nullptr && n > 0, "fc2td scratch: no device address");
    fc2td_ncu[dev] = n;
    fc2td_scr[dev] = static_cast<double*>(p);
  }
  *out = fc2td_scr[dev];
  *ncu = fc2td_ncu[dev];
  return DOGPUP_OK;
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

int dogpup_fc2td_field(const double* d_field, float* d_out, const double* d_f, const double* d_t, int nf, int nt,
                       long long ncol, void* stream) {
  DP_ARG(d_field && d_out && d_f && d_t && nf >= 1 && nf <= 128 && nt >= 1 && ncol >= 1, "bad arguments");
  hipStream_t st = as_stream(stream);
  int dev = 0, ncu = 0;
  DP_HIP(hipGetDevice(&dev));
  DP_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const Fc2tdShape sh = fc2td_shape(nf);
  double* Wd = nullptr;
  DP_HIP(hipMalloc((void**)&Wd, (size_t)sh.bins() * sh.kks * sizeof(double)));
  const hipError_t e = fc2td_field_passes(d_field, d_out, d_f, d_t, nf, nt, ncol, Wd, ncu, st);
  const hipError_t e2 = hipStreamSynchronize(st);   // Wd is freed after the kernels used it
  (void)hipFree(Wd);
  if (e != hipSuccess) return fail(DOGPUP_ERR_HIP, "fc2td_field launch: %s", hipGetErrorString(e));
  if (e2 != hipSuccess) return fail(DOGPUP_ERR_HIP, "fc2td_field: %s", hipGetErrorString(e2));
  return DOGPUP_OK;
}

// Allocation-free fc2td_field: the per-pass DFT table lives in `scratch`, a
// static device array (one copy per GPU, kept until the process ends).  Host
// code never passes the array's name to a kernel: its device address is
// resolved once per device with hipGetSymbolAddress and kept in fc2td_scr.
int dogpup_fc2td_field_noalloc(const double* d_field, float* d_out, const double* d_f, const double* d_t, int nf,
                               int nt, long long ncol, void* stream) {
  DP_ARG(d_field && d_out && d_f && d_t && nf >= 1 && nf <= 128 && nt >= 1 && ncol >= 1, "bad arguments");
  const Fc2tdShape sh = fc2td_shape(nf);
  const long long need = (long long)sh.bins() * sh.kks;
  DP_ARG(need <= kFc2tdScratch, "fc2td_field_noalloc: nf = %d needs %lld doubles of scratch, the static buffer holds %d",
         nf, need, kFc2tdScratch);
  double* wd = nullptr;
  int ncu = 0;
  const int rc = fc2td_scratch(&wd, &ncu);
  if (rc != DOGPUP_OK) return rc;
  const hipError_t e = fc2td_field_passes(d_field, d_out, d_f, d_t, nf, nt, ncol, wd, ncu, as_stream(stream));
  if (e != hipSuccess) return fail(DOGPUP_ERR_HIP, "fc2td_field_noalloc launch: %s", hipGetErrorString(e));
  return DOGPUP_OK;
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
  hipStream_t st = as_stream(stream);
  // MFMA form when the links span at most 32 source and 32 detector columns
  // without repeated pairs (the table is built from a host copy of link)
  const char* force_scalar = getenv("DOGPUP_JTG_SCALAR");
  if (nf <= 128 && !(force_scalar && atoi(force_scalar) != 0)) {
    std::vector<int> hl((size_t)2 * nm);
    DP_HIP(hipMemcpyAsync(hl.data(), d_link, hl.size() * 4, hipMemcpyDeviceToHost, st));
    DP_HIP(hipStreamSynchronize(st));
    std::vector<int> tab(64 + 1024, -1);
    int nd = 0, ns = 0;
    bool ok = true;
    auto slot = [&](int* cols, int& n, int c) -> int {
      for (int i = 0; i < n; ++i)
        if (cols[i] == c) return i;
      if (n == 32) return -1;
      cols[n] = c;
      return n++;
    };
    for (int m = 0; m < nm && ok; ++m) {
      const int s = hl[2 * m], d = hl[2 * m + 1];
      DP_ARG(s >= 0 && s < ncols && d >= 0 && d < ncols, "link %d names a column outside the field", m);
      const int j = slot(tab.data() + 32, ns, s), i = slot(tab.data(), nd, d);
      if (i < 0 || j < 0 || tab[64 + i * 32 + j] >= 0) ok = false;
      else tab[64 + i * 32 + j] = m;
    }
    if (ok) {
      int* d_tab = nullptr;
      DP_HIP(hipMalloc((void**)&d_tab, tab.size() * 4));
      DP_HIP(hipMemcpyAsync(d_tab, tab.data(), tab.size() * 4, hipMemcpyHostToDevice, st));
      for (int gb = 0; gb < ng; gb += 4) {
        k_jtg_mfma<4><<<div_up(nv, 2), 256, 0, st>>>(d_field, d_coef, d_tab, d_scale, d_out, nn, nf, ng, gb, ncols,
                                                     nsp, nm, nv);
        DP_LAUNCH_CHECK();
      }
      DP_HIP(hipStreamSynchronize(st));   // the table is freed after the kernels used it
      DP_HIP(hipFree(d_tab));
      return DOGPUP_OK;
    }
  }
  DP_ARG(ng <= 16, "jacobian_tg supports at most 16 gates");
  // one node's fluence and the gate coefficients in LDS as complex128 (up to
  // the 160 KiB of a CU)
  const size_t shm = (size_t)(nf * ncols + ng * nf) * sizeof(double2);
  DP_ARG(shm <= 160 * 1024, "nf*(ncols+ng) too large for one node's LDS stage (%zu bytes)", shm);
  if (ng <= 8) {
    DP_HIP(hipFuncSetAttribute((const void*)k_jacobian_tg<8>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    k_jacobian_tg<8><<<nv, 256, shm, st>>>(d_field, d_coef, d_link, d_scale, d_out, nn, nf, ng, ncols,
                                           nsp, nm);
  } else {
    DP_HIP(hipFuncSetAttribute((const void*)k_jacobian_tg<16>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)shm));
    k_jacobian_tg<16><<<nv, 256, shm, st>>>(d_field, d_coef, d_link, d_scale, d_out, nn, nf, ng, ncols,
                                            nsp, nm);
  }
  DP_LAUNCH_CHECK();
  return DOGPUP_OK;
}

}  // extern "C"
