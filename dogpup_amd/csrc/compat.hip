// Reference-compatible entry points: the launchers smv_call / FSAImv_call /
// FSAIP_call (planar FP64, reference layout, device pointers) and the
// mex-level calls smv_mex / FSAImv_mex / FSAIP_gen_mex (host interleaved
// complex arrays).
//
// Reference layout: x(row, col, freq) at (freq*ncols + col)*nrows + row, so a
// wavefront of consecutive rows reads consecutive addresses of every column.
// Each lane owns one row and CB columns of one frequency and walks its CSR
// row once, reusing every matrix entry for its CB columns.
#include <cstdarg>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "fsai.hpp"

namespace dogpup {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

// Access to a complex element in planar (re, im arrays) or interleaved form.
struct CRef {
  const double* re;
  const double* im;  // nullptr = interleaved in re
  __device__ inline double r(long long i) const { return im ? re[i] : re[2 * i]; }
  __device__ inline double ii(long long i) const { return im ? im[i] : re[2 * i + 1]; }
};
struct WRef {
  double* re;
  double* im;
  __device__ inline void put(long long i, double a, double b) const {
    if (im) { re[i] = a; im[i] = b; }
    else { re[2 * i] = a; re[2 * i + 1] = b; }
  }
};

constexpr int CB = 4;  // columns per lane

// MAT = 0: A shared by all frequencies, Im scaled by fAxis[freq].
// MAT = 1: per-frequency values at valG[nnz*freq + i].
template <int MAT>
__global__ __launch_bounds__(256) void k_compat_csrmv(WRef y, CRef x, CRef a, const double* __restrict__ fAxis,
                                                      const int* __restrict__ rptr,
                                                      const int* __restrict__ cptr, int nnz,
                                                      int nrows, int ncols, int nfreqs) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  const int c0 = blockIdx.y * CB;
  const int freq = blockIdx.z;
  if (row >= nrows) return;
  const double fk = (MAT == 0) ? fAxis[freq] : 0.0;
  const long long aoff = (MAT == 1) ? (long long)nnz * freq : 0;
  double accr[CB], acci[CB];
#pragma unroll
  for (int c = 0; c < CB; ++c) { accr[c] = 0.0; acci[c] = 0.0; }
  const int s0 = rptr[row], s1 = rptr[row + 1];
  for (int p = s0; p < s1; ++p) {
    const int col = cptr[p];
    const double ar = a.r(aoff + p);
    const double ai = (MAT == 0) ? fk * a.ii(p) : a.ii(aoff + p);
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      if (c0 + c < ncols) {
        const long long xi = ((long long)freq * ncols + c0 + c) * nrows + col;
        const double xr = x.r(xi), xm = x.ii(xi);
        accr[c] += ar * xr - ai * xm;
        acci[c] += ar * xm + ai * xr;
      }
    }
  }
#pragma unroll
  for (int c = 0; c < CB; ++c)
    if (c0 + c < ncols) y.put(((long long)freq * ncols + c0 + c) * nrows + row, accr[c], acci[c]);
}

template <int MAT>
static int compat_csrmv(WRef y, CRef x, CRef a, const double* fAxis, const int* rptr,
                        const int* cptr, int nnz, int nrows, int ncols, int nfreqs,
                        hipStream_t st) {
  DP_ARG(nrows >= 0 && ncols >= 0 && nfreqs >= 0, "negative dimension");
  if (nrows == 0 || ncols == 0 || nfreqs == 0) return DOGPUP_OK;
  DP_ARG(nfreqs <= 65535 && div_up(ncols, CB) <= 65535, "too many columns/frequencies");
  dim3 grid(div_up(nrows, 256), div_up(ncols, CB), nfreqs);
  k_compat_csrmv<MAT><<<grid, 256, 0, st>>>(y, x, a, fAxis, rptr, cptr, nnz, nrows, ncols, nfreqs);
  DP_LAUNCH_CHECK();
  return DOGPUP_OK;
}

// Host-side RAII buffer for the mex-level calls.
struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t bytes) {
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(DOGPUP_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return DOGPUP_OK;
  }
  int upload(const void* h, size_t bytes) {
    int rc = alloc(bytes);
    if (rc) return rc;
    if (bytes) DP_HIP(hipMemcpy(p, h, bytes, hipMemcpyHostToDevice));
    return DOGPUP_OK;
  }
  template <typename T>
  T* as() const { return reinterpret_cast<T*>(p); }
};

#define DP_TRY(x)              \
  do {                         \
    int rc_ = (x);             \
    if (rc_ != DOGPUP_OK) return rc_; \
  } while (0)

}  // namespace dogpup

using namespace dogpup;

extern "C" {

const char* dogpup_last_error(void) { return g_err.c_str(); }

int dogpup_version(void) { return 1; }

int dogpup_device_count(int* count) {
  DP_ARG(count != nullptr, "count is NULL");
  DP_HIP(hipGetDeviceCount(count));
  return DOGPUP_OK;
}

int dogpup_smv_call(double* d_out_real, double* d_out_imag, const double* d_x_real,
                    const double* d_x_imag, const double* d_fAxis, const double* d_valA_real,
                    const double* d_valA_imag, const int* d_rPtr, const int* d_cPtr, int nrows,
                    int ncols, int nfreqs, void* stream) {
  DP_ARG(d_out_real && d_out_imag && d_x_real && d_x_imag && d_fAxis && d_valA_real &&
             d_valA_imag && d_rPtr && d_cPtr,
         "smv_call: NULL pointer");
  return compat_csrmv<0>(WRef{d_out_real, d_out_imag}, CRef{d_x_real, d_x_imag},
                         CRef{d_valA_real, d_valA_imag}, d_fAxis, d_rPtr, d_cPtr, 0, nrows, ncols,
                         nfreqs, as_stream(stream));
}

int dogpup_fsaimv_call(double* d_out_real, double* d_out_imag, const double* d_x_real,
                       const double* d_x_imag, const double* d_valG_real,
                       const double* d_valG_imag, const int* d_rPtr, const int* d_cPtr, int nnz,
                       int nrows, int ncols, int nfreqs, void* stream) {
  DP_ARG(d_out_real && d_out_imag && d_x_real && d_x_imag && d_valG_real && d_valG_imag &&
             d_rPtr && d_cPtr,
         "FSAImv_call: NULL pointer");
  return compat_csrmv<1>(WRef{d_out_real, d_out_imag}, CRef{d_x_real, d_x_imag},
                         CRef{d_valG_real, d_valG_imag}, nullptr, d_rPtr, d_cPtr, nnz, nrows,
                         ncols, nfreqs, as_stream(stream));
}

int dogpup_fsaip_call(double* d_out_real, double* d_out_imag, const double* d_valA_real,
                      const double* d_valA_imag, const int* d_rPtrA, const int* d_cPtrA,
                      const int* d_rPtrG, const int* d_cPtrG, int nfreqs, int nrows,
                      void* stream) {
  DP_ARG(d_out_real && d_out_imag && d_valA_real && d_valA_imag && d_rPtrA && d_cPtrA &&
             d_rPtrG && d_cPtrG,
         "FSAIP_call: NULL pointer");
  DP_ARG(nrows >= 0 && nfreqs >= 0, "negative dimension");
  FsaiOut out{d_out_real, d_out_imag, 0, FSAI_OUT_REF};
  return fsai_launch(d_valA_real, d_valA_imag, 0, d_rPtrA, d_cPtrA, d_rPtrG, d_cPtrG, nrows,
                     nfreqs, nullptr, out, as_stream(stream));
}

// Host check of a CSR structure that a mex-level call receives as host
// arrays: the row pointer starts at 0, never decreases and ends at nnz, and
// every column lies in [0, ncol).  A malformed structure would send the
// kernels out of bounds, so it is an argument error before any upload.
static int check_csr_host(const char* who, const char* name, const int* rp, const int* cp, int nrows,
                          int nnz, int ncol) {
  DP_ARG(rp[0] == 0 && rp[nrows] == nnz, "%s: rPtr%s must run from 0 to nnz = %d (got %d .. %d)", who,
         name, nnz, rp[0], rp[nrows]);
  for (int i = 0; i < nrows; ++i)
    DP_ARG(rp[i] <= rp[i + 1], "%s: rPtr%s decreases at row %d", who, name, i);
  for (int q = 0; q < nnz; ++q)
    DP_ARG(cp[q] >= 0 && cp[q] < ncol, "%s: cPtr%s[%d] = %d outside [0, %d)", who, name, q, cp[q], ncol);
  return DOGPUP_OK;
}

int dogpup_smv_mex(double* h_y, const double* h_x, const double* h_f, const double* h_A,
                   const int* h_rPtrA, const int* h_cPtrA, int nrows, int ncols, int nfreqs,
                   int nnz) {
  DP_ARG(h_y && h_x && h_f && h_A && h_rPtrA && h_cPtrA, "smv_mex: NULL pointer");
  DP_ARG(nrows >= 0 && ncols >= 0 && nfreqs >= 0 && nnz >= 0, "smv_mex: negative dimension");
  DP_TRY(check_csr_host("smv_mex", "A", h_rPtrA, h_cPtrA, nrows, nnz, nrows));
  const size_t nel = (size_t)nrows * ncols * nfreqs;
  DevBuf x, f, a, rp, cp, y;
  DP_TRY(x.upload(h_x, nel * 16));
  DP_TRY(f.upload(h_f, (size_t)nfreqs * 8));
  DP_TRY(a.upload(h_A, (size_t)nnz * 16));
  DP_TRY(rp.upload(h_rPtrA, (size_t)(nrows + 1) * 4));
  DP_TRY(cp.upload(h_cPtrA, (size_t)nnz * 4));
  DP_TRY(y.alloc(nel * 16));
  DP_TRY(compat_csrmv<0>(WRef{y.as<double>(), nullptr}, CRef{x.as<double>(), nullptr},
                         CRef{a.as<double>(), nullptr}, f.as<double>(), rp.as<int>(), cp.as<int>(),
                         0, nrows, ncols, nfreqs, 0));
  DP_HIP(hipMemcpy(h_y, y.p, nel * 16, hipMemcpyDeviceToHost));
  return DOGPUP_OK;
}

int dogpup_fsaimv_mex(double* h_y, const double* h_x, const double* h_G, const int* h_rPtrG,
                      const int* h_cPtrG, int nrows, int ncols, int nfreqs, int nnzG) {
  DP_ARG(h_y && h_x && h_G && h_rPtrG && h_cPtrG, "FSAImv_mex: NULL pointer");
  DP_ARG(nrows >= 0 && ncols >= 0 && nfreqs >= 0 && nnzG >= 0, "FSAImv_mex: negative dimension");
  DP_TRY(check_csr_host("FSAImv_mex", "G", h_rPtrG, h_cPtrG, nrows, nnzG, nrows));
  const size_t nel = (size_t)nrows * ncols * nfreqs;
  DevBuf x, g, rp, cp, y;
  DP_TRY(x.upload(h_x, nel * 16));
  DP_TRY(g.upload(h_G, (size_t)nnzG * nfreqs * 16));
  DP_TRY(rp.upload(h_rPtrG, (size_t)(nrows + 1) * 4));
  DP_TRY(cp.upload(h_cPtrG, (size_t)nnzG * 4));
  DP_TRY(y.alloc(nel * 16));
  DP_TRY(compat_csrmv<1>(WRef{y.as<double>(), nullptr}, CRef{x.as<double>(), nullptr},
                         CRef{g.as<double>(), nullptr}, nullptr, rp.as<int>(), cp.as<int>(), nnzG,
                         nrows, ncols, nfreqs, 0));
  DP_HIP(hipMemcpy(h_y, y.p, nel * 16, hipMemcpyDeviceToHost));
  return DOGPUP_OK;
}

int dogpup_fsaip_gen_mex(double* h_valG, const double* h_valA, const int* h_rPtrA,
                         const int* h_cPtrA, const int* h_rPtrG, const int* h_cPtrG, int nrows,
                         int nnzA, int nnzG, int nfreqs) {
  DP_ARG(h_valG && h_valA && h_rPtrA && h_cPtrA && h_rPtrG && h_cPtrG,
         "FSAIP_gen_mex: NULL pointer");
  DP_ARG(nrows >= 0 && nnzA >= 0 && nnzG >= 0 && nfreqs >= 0, "FSAIP_gen_mex: negative dimension");
  DP_TRY(check_csr_host("FSAIP_gen_mex", "A", h_rPtrA, h_cPtrA, nrows, nnzA, nrows));
  DP_TRY(check_csr_host("FSAIP_gen_mex", "G", h_rPtrG, h_cPtrG, nrows, nnzG, nrows));
  DevBuf a, ra, ca, rg, cg, outr, outi;
  DP_TRY(a.upload(h_valA, (size_t)nnzA * 16));
  DP_TRY(ra.upload(h_rPtrA, (size_t)(nrows + 1) * 4));
  DP_TRY(ca.upload(h_cPtrA, (size_t)nnzA * 4));
  DP_TRY(rg.upload(h_rPtrG, (size_t)(nrows + 1) * 4));
  DP_TRY(cg.upload(h_cPtrG, (size_t)nnzG * 4));
  const size_t nout = (size_t)nnzG * nfreqs;
  DP_TRY(outr.alloc(nout * 8));
  DP_TRY(outi.alloc(nout * 8));
  FsaiOut out{outr.as<double>(), outi.as<double>(), nnzG, FSAI_OUT_REF};
  DP_TRY(fsai_launch(a.as<double>(), nullptr, 1, ra.as<int>(), ca.as<int>(), rg.as<int>(),
                     cg.as<int>(), nrows, nfreqs, nullptr, out, 0));
  std::vector<double> re(nout), im(nout);
  DP_HIP(hipMemcpy(re.data(), outr.p, nout * 8, hipMemcpyDeviceToHost));
  DP_HIP(hipMemcpy(im.data(), outi.p, nout * 8, hipMemcpyDeviceToHost));
  // NF x nnzG column-major = frequency fastest, interleaved complex
  for (size_t i = 0; i < nout; ++i) {
    h_valG[2 * i] = re[i];
    h_valG[2 * i + 1] = im[i];
  }
  return DOGPUP_OK;
}

}  // extern "C"
